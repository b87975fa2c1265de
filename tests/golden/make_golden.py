"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in a container that holds the reference sources (oracle/Makefile builds
oracle/_ref/libxfref.so from them):  python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  ref_beam12.npz   12x4x4 beam (0.1 m cells, x from 0): the reference's
                   element constants (Qi, QQ, QR, volume, w), its MINSTD
                   tOrder, GenerateTetBlock output for a centred 12x4x4 block,
                   the greedy natural-order colour sweep order, and node
                   positions/velocities of GeoLinear3d::Substep (x = 0 face
                   pinned, rho 1000, compliance 1e-5, nu 0.5, g = -9.81,
                   dt = 1/1200) at frames 1, 10 and 120 (20 substeps each)
                   for MixedSel and YeohSkinFast (simultaneous) and MixedSel
                   serial.
  ref_tet_kats.npz 64 random deformed tets: inputs (X, w, element constants,
                   energy, solve mode) and the reference SolveElement output.
  ref_passes.npz   6x3x3 block (0.1 m cells, x = 0 face pinned, MINSTD
                   tOrder): node positions/velocities after 20 substeps of
                   GeoLinear3d::Substep with volume passes and with each
                   Rayleigh damping type plus PBD damping (tickId from 5),
                   for the cases listed in its `cases` array
                   (energy, simultaneous, volume passes, rayleigh type,
                   damping, pbd factor, amortized pbd factor).
  ref_rings.npz    the same 6x3x3 block with the vertex-ring energies
                   (ContinuousPixar 8, ContinuousMixed 9, ContinuousSkin 10):
                   the reference's OneRing arrays (CSR: ring_node_off,
                   ring_nodes, ring_tet_off, ring_tets, ring_idx), its MINSTD
                   vOrder and tOrder, and node positions/velocities after 20
                   substeps (tickId from 5) for the cases of its `cases` array
                   (energy, Poisson ratio, compliance, rayleigh type, damping,
                   pbd factor, amortized pbd factor).
  ref_dispatch.npz the 6x3x3 block (x = 0 face pinned, MINSTD tOrder, tickId
                   from 5): positions/velocities after 40 substeps for the rest
                   of TSolveElement's switch (Pixar 0/1/2 with every Rayleigh
                   type and PBD, Cube 11/12, defaults 13/31), per its `cases`
                   array (energy, simultaneous, nu, compliance, rayleigh type,
                   damping, pbd factor, amortized pbd factor).
  ref_autoresize.npz
                   the 12x4x4 beam through GeoLinear3d::Init with autoResize:
                   the FP64 positions Init computes, and positions/velocities
                   after 40 substeps (MixedSel, simultaneous, compliance 1e-5,
                   nu 0.5, x = 0 face pinned, dt = 1/1200) over the greedy
                   natural-order colour sweep order.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle")]

import pyoracle as po  # noqa: E402

FRAMES = (1, 10, 120)
RUNS = ((4, True), (7, True), (4, False))
DT = np.float32(1.0) / np.float32(1200.0)


def natural_colour_order(n, tets):
    t = tets.size // 4
    nc, col = po.colour_greedy(n, tets, np.arange(t, dtype=np.uint32))
    return np.concatenate([np.nonzero(col == c)[0] for c in range(nc)]).astype(np.uint32)


def beam12():
    xyz, tets = po.tet_block(12, 4, 4, (0.1, 0.1, 0.1), (0.0, -0.2, -0.2))
    n, t = xyz.size // 3, tets.size // 4
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    out = dict(xyz=xyz, tets=tets, pinned=pinned, dt=np.array([DT], np.float32))
    rb = po.RefBody(xyz, tets, 1000.0, None)
    el = rb.elements()
    for k, v in el.items():
        out["el_" + k] = v
    out["minstd_order"] = rb.order()
    cx = np.zeros(3 * 13 * 5 * 5, np.float32)
    ct = np.zeros(4 * t, np.uint32)
    po.ref().xr_gen_tet_block(12, 4, 4, 0.1, 0.1, 0.1, 0, cx.ctypes.data, ct.ctypes.data)
    out["gen_xyz"], out["gen_tets"] = cx, ct
    order = natural_colour_order(n, tets)
    out["sweep_order"] = order
    for energy, sim in RUNS:
        tag = f"e{energy}{'s' if sim else 'g'}"
        rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
        st = rb.settings(1e-5, nu=0.5, energy=energy, simultaneous=sim)
        done = 0
        for f in FRAMES:
            rb.substeps(st, DT, 20 * f - done)
            done = 20 * f
            X, V = rb.state()
            out[f"{tag}_X{f}"] = X
            out[f"{tag}_V{f}"] = V
    np.savez_compressed(os.path.join(HERE, "ref_beam12.npz"), **out)


def tet_kats(count=64, seed=7):
    rng = np.random.default_rng(seed)
    rows = dict(X=[], w=[], qi=[], qq=[], qr=[], vol=[], energy=[], sim=[], Xout=[])
    for k in range(count):
        # random well-shaped rest tet, then a random deformation of it
        rest = (np.eye(3, dtype=np.float64) * 0.1 + rng.normal(0, 0.02, (3, 3))).astype(np.float32)
        if np.linalg.det(rest) < 0:
            rest[0] *= -1
        xyz = np.vstack([rest, np.zeros((1, 3), np.float32)]).astype(np.float32).reshape(-1)
        el = po.init_elements(xyz, np.arange(4, dtype=np.uint32), 1000.0)
        F = np.eye(3) + rng.normal(0, 0.05, (3, 3))
        X = (xyz.reshape(4, 3).astype(np.float64) @ F.T + rng.normal(0, 1e-3, 3)).reshape(-1)
        w = (el["w"] * rng.uniform(0.5, 2.0, 4)).astype(np.float32)
        if k % 8 == 0:
            w[rng.integers(0, 4)] = 0.0
        energy = [3, 4, 5, 6, 7][k % 5]
        sim = (k // 5) % 2 == 0
        st = po.XrSettings()
        st.compliance = 1e-5
        st.poissons_ratio = 0.5 if k % 3 else 0.45
        st.gravity[:] = (0.0, -9.81)
        st.drag = 0.0
        st.energy = energy
        st.simultaneous = 1 if sim else 0
        Xo = X.copy()
        po.ref().xr_solve_tet(st, DT, Xo, w, el["qi"], el["qq"], el["qr"], float(el["vol"][0]))
        for key, val in (("X", X), ("w", w), ("qi", el["qi"]), ("qq", el["qq"]), ("qr", el["qr"]),
                         ("vol", el["vol"]), ("energy", energy), ("sim", sim), ("Xout", Xo)):
            rows[key].append(val)
        rows.setdefault("nu", []).append(st.poissons_ratio)
    out = {k: np.asarray(v) for k, v in rows.items()}
    out["dt"] = np.array([DT], np.float32)
    out["compliance"] = np.array([1e-5], np.float32)
    np.savez_compressed(os.path.join(HERE, "ref_tet_kats.npz"), **out)


PASS_CASES = (
    # energy, simultaneous, volume passes, rayleigh type, damping, pbd, pbd amortized
    (4, 1, 2, 0, 0.0, 0.0, 0.0),
    (7, 0, 1, 0, 0.0, 0.0, 0.0),
    (4, 1, 0, 0, 0.02, 0.0, 0.0),
    (4, 0, 0, 1, 0.02, 0.0, 0.0),
    (7, 1, 0, 1, 0.02, 0.0, 0.0),
    (4, 1, 0, 2, 0.02, 0.003, 0.012),
    (3, 1, 0, 3, 0.02, 0.003, 0.012),
    (7, 0, 0, 3, 0.02, 0.003, 0.012),
    (4, 1, 1, 3, 0.05, 0.003, 0.012),
)


def passes():
    xyz, tets = po.tet_block(6, 3, 3, (0.1, 0.1, 0.1), (0.0, -0.15, -0.15))
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    order = po.minstd_order(tets.size // 4)
    out = dict(xyz=xyz, tets=tets, pinned=pinned, order=order, dt=np.array([DT], np.float32),
               compliance=np.array([2e-5], np.float32), tick0=np.array([5], np.uint32),
               cases=np.array(PASS_CASES, np.float64))
    for k, (energy, sim, vp, rtype, damp, pbd, pbda) in enumerate(PASS_CASES):
        rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
        st = rb.settings(2e-5, energy=int(energy), simultaneous=bool(sim), volume_passes=int(vp), damping=damp,
                         rayleigh_type=int(rtype), pbd_damping=1.0 if pbd else 0.0, volume_pbd=pbd,
                         amortized_volume_pbd=pbda, tick_id=5)
        rb.substeps(st, DT, 20)
        out[f"X{k}"], out[f"V{k}"] = rb.state()
    np.savez_compressed(os.path.join(HERE, "ref_passes.npz"), **out)


RING_CASES = (
    # energy, nu, compliance, rayleigh type, damping, pbd, pbd amortized
    (8, 0.45, 1e-6, 0, 0.0, 0.0, 0.0),
    (8, 0.3, 1e-6, 0, 0.02, 0.0, 0.0),
    (8, 0.45, 1e-6, 3, 0.02, 0.003, 0.012),
    (9, 0.45, 2e-5, 0, 0.0, 0.0, 0.0),
    (9, 0.5, 2e-5, 1, 0.02, 0.0, 0.0),
    (9, 0.45, 2e-5, 2, 0.02, 0.003, 0.012),
    (10, 0.45, 2e-5, 0, 0.0, 0.0, 0.0),
    (10, 0.45, 2e-5, 0, 0.02, 0.0, 0.0),
    (10, 0.4, 2e-5, 2, 0.05, 0.0, 0.0),
    (10, 0.45, 2e-5, 3, 0.02, 0.003, 0.012),
)


def rings():
    xyz, tets = po.tet_block(6, 3, 3, (0.1, 0.1, 0.1), (0.0, -0.15, -0.15))
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    rb = po.RefBody(xyz, tets, 1000.0, pinned)
    n_rings = rb.ring_order().size
    node_off, nodes, tet_off, rtets, idx = [0], [], [0], [], []
    for v in range(n_rings):
        a, b, c = rb.ring(v)
        nodes.append(a)
        rtets.append(b)
        idx.append(c.reshape(-1))
        node_off.append(node_off[-1] + a.size)
        tet_off.append(tet_off[-1] + b.size)
    out = dict(xyz=xyz, tets=tets, pinned=pinned, ring_order=rb.ring_order(), order=rb.order(),
               dt=np.array([DT], np.float32), tick0=np.array([5], np.uint32), cases=np.array(RING_CASES, np.float64),
               ring_node_off=np.array(node_off, np.uint32), ring_nodes=np.concatenate(nodes).astype(np.uint32),
               ring_tet_off=np.array(tet_off, np.uint32), ring_tets=np.concatenate(rtets).astype(np.uint32),
               ring_idx=np.concatenate(idx).astype(np.uint8))
    for k, (energy, nu, comp, rtype, damp, pbd, pbda) in enumerate(RING_CASES):
        rb = po.RefBody(xyz, tets, 1000.0, pinned)
        st = rb.settings(comp, nu=nu, energy=int(energy), damping=damp, rayleigh_type=int(rtype),
                         pbd_damping=1.0 if pbd else 0.0, volume_pbd=pbd, amortized_volume_pbd=pbda, tick_id=5)
        rb.substeps(st, DT, 20)
        out[f"X{k}"], out[f"V{k}"] = rb.state()
    np.savez_compressed(os.path.join(HERE, "ref_rings.npz"), **out)


DISPATCH_CASES = (
    # energy, simultaneous, nu, compliance, rayleigh type, damping, pbd, pbd amortized
    (0, 1, 0.45, 1e-6, 0, 0.0, 0.0, 0.0),
    (1, 0, 0.3, 1e-6, 0, 0.0, 0.0, 0.0),
    (2, 1, 0.45, 1e-6, 0, 0.0, 0.0, 0.0),
    (0, 1, 0.45, 1e-6, 1, 0.02, 0.0, 0.0),
    (2, 1, 0.45, 1e-6, 0, 0.02, 0.0, 0.0),
    (0, 1, 0.45, 1e-6, 2, 0.02, 0.003, 0.0),
    (2, 0, 0.45, 1e-6, 3, 0.02, 0.003, 0.012),
    (11, 1, 0.45, 2e-5, 0, 0.0, 0.0, 0.0),
    (12, 1, 0.45, 2e-5, 3, 0.02, 0.003, 0.012),
    (13, 1, 0.45, 2e-5, 0, 0.0, 0.0, 0.0),
    (31, 0, 0.45, 2e-5, 2, 0.02, 0.0, 0.0),
)


def dispatch():
    xyz, tets = po.tet_block(6, 3, 3, (0.1, 0.1, 0.1), (0.0, -0.15, -0.15))
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    order = po.minstd_order(tets.size // 4)
    out = dict(xyz=xyz, tets=tets, pinned=pinned, order=order, cases=np.array(DISPATCH_CASES, np.float64))
    for k, (energy, sim, nu, comp, rtype, damp, pbd, pbda) in enumerate(DISPATCH_CASES):
        rb = po.RefBody(xyz, tets, 1000.0, pinned, order)
        st = rb.settings(comp, nu=nu, energy=energy, simultaneous=bool(sim), damping=damp, rayleigh_type=rtype,
                         pbd_damping=1.0 if pbd else 0.0, volume_pbd=pbd, amortized_volume_pbd=pbda, tick_id=5)
        rb.substeps(st, DT, 40)
        out[f"X{k}"], out[f"V{k}"] = rb.state()
    np.savez_compressed(os.path.join(HERE, "ref_dispatch.npz"), **out)


def autoresize():
    xyz, tets = po.tet_block(12, 4, 4, (0.1, 0.1, 0.1), (0.0, -0.2, -0.2))
    pinned = (xyz[0::3] == 0.0).astype(np.uint8)
    n = xyz.size // 3
    order = natural_colour_order(n, tets)
    rb = po.RefBody(xyz, tets, 1000.0, pinned, order, autoresize=True)
    X0, _ = rb.state()
    rb.substeps(rb.settings(1e-5, nu=0.5, energy=4, simultaneous=True), DT, 40)
    X, V = rb.state()
    np.savez_compressed(os.path.join(HERE, "ref_autoresize.npz"), xyz=xyz, tets=tets, pinned=pinned, order=order,
                        X0=X0, X=X, V=V)


if __name__ == "__main__":
    if not po.have_ref():
        sys.exit("oracle/_ref/libxfref.so is missing: build the oracle with the reference present")
    beam12()
    tet_kats()
    passes()
    rings()
    autoresize()
    dispatch()
    for f in ("ref_beam12.npz", "ref_tet_kats.npz", "ref_passes.npz", "ref_rings.npz", "ref_autoresize.npz",
              "ref_dispatch.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
