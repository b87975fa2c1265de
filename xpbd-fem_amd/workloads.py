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


def beam_yeoh_96x24x24(seed=1, h=0.01, structured=True):
    """96x24x24 cells of h = 0.01 m (331,776 tets, 60,625 nodes), Yeoh
    C10 = 5e4, C20 = 5e3, C30 = 5e2 Pa (compliance 1/C10, C = (1, 0.1, 0.01),
    YeohSkinFast form), incompressible, density 1000, initial velocity
    N(0, 0.05^2) per component from SplitMix64(seed) + Box-Muller, x = 0 face
    pinned, ground y = -0.5 (friction 0.5) and a kinematic sphere r = 0.3 m
    starting at (0, -0.40, 0) moving +x at 1 m/s; 1/60 s frames of 30
    substeps, 60 frames."""
    W, H, D = 96, 24, 24
    xyz, tets = xf.tet_block(W, H, D, (h, h, h), (0.0, -0.5 * H * h, -0.5 * D * h))
    n = xyz.size // 3
    vel = xf.normal_fill(seed, 0.05, 3 * n)
    substeps = 30
    c10, c20, c30 = 5e4, 5e3, 5e2
    return dict(
        name="beam96x24x24_yeoh", xyz=xyz, tets=tets, density=1000.0, pinned=_pins_x0(xyz), velocity=vel,
        tet_mu=None, tet_colour=xf.tet_block_colours(W, H, D) if structured else None,
        settings=dict(energy=xf.ENERGY_YEOH_SKIN_FAST, simultaneous=1, compliance=1.0 / c10, poissons_ratio=0.5,
                      gravity=(0.0, -9.81), yeoh=(1.0, c20 / c10, c30 / c10),
                      collide=xf.COLLIDE_GROUND | xf.COLLIDE_SPHERE, ground_y=-0.5, friction=0.5,
                      sphere_c0=(0.0, -0.40, 0.0), sphere_v=(1.0, 0.0, 0.0), sphere_r=0.3),
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


def make_batch_scene(ctx, batch, bodies=None, colour_order=xf.COLOUR_NATURAL, use_graph=True, block_size=64,
                     launch_mode=xf.LAUNCH_PER_COLOUR, persistent_blocks=0):
    """One scene holding the given bodies of a batch (all by default)."""
    sc = xf.Scene(ctx)
    chosen = batch["bodies"] if bodies is None else [batch["bodies"][i] for i in bodies]
    for wl in chosen:
        sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"], velocity=wl["velocity"],
                    tet_mu=wl["tet_mu"], tet_colour=wl.get("tet_colour"))
    sc.finalize(colour_order=colour_order, use_graph=use_graph, block_size=block_size, launch_mode=launch_mode,
                persistent_blocks=persistent_blocks)
    for i, wl in enumerate(chosen):
        sc.set_settings(i, **wl["settings"])
    return sc


def make_scene(ctx, wl, colour_order=xf.COLOUR_NATURAL, use_graph=True, block_size=64,
               launch_mode=xf.LAUNCH_PER_COLOUR, persistent_blocks=0):
    """Builds and finalizes a one-body scene of a workload."""
    sc = xf.Scene(ctx)
    sc.add_body(wl["xyz"], wl["tets"], wl["density"], pinned=wl["pinned"], velocity=wl["velocity"],
                tet_mu=wl["tet_mu"], tet_colour=wl.get("tet_colour"))
    sc.finalize(colour_order=colour_order, use_graph=use_graph, block_size=block_size, launch_mode=launch_mode,
                persistent_blocks=persistent_blocks)
    sc.set_settings(0, **wl["settings"])
    return sc
