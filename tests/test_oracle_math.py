"""Closed-form checks of the oracle (and of the host utilities of the
product), independent of the reference's own code.

* mesh counts, rest volumes and total lumped mass;
* Qi * Ds^T = I;
* the simultaneous projection against a textbook restatement:
  F = Ds Dm^-1, U0 = |F|_F^2 (MixedSel), U0 = C0 IM + C1 IM^2 + C2 IM^3
  (Yeoh), U1 = (det F - a)^2, analytic gradients, A lambda = -2U solved
  with numpy (catches a dropped term, a wrong sign or a transposed operand);
* linear momentum and angular momentum of a projection are conserved;
* energies 3 and 5 equal 4 and 7 for a linear tet (one quadrature point);
* collision pass: non-penetration, static and kinetic friction on the ground
  and on the moving sphere (Coulomb geometry at many contact normals, each
  check shown to reject planted mistakes);
* colouring: node-disjoint colours, product == oracle;
* LDLT usage example of Ldlt.h: x = (-3, 3, 1).
"""
import numpy as np
import pytest

import pyoracle as po
import xpbdfem as xf

DT = float(np.float32(1.0) / np.float32(1200.0))


def rest_tet(rng):
    rest = np.eye(3) * 0.1 + rng.normal(0, 0.02, (3, 3))
    if np.linalg.det(rest) < 0:
        rest[0] *= -1
    return np.vstack([rest, np.zeros((1, 3))]).astype(np.float32)


def textbook(X, rest, w, energy, comp0, comp1, a, C=po.SKIN_YEOH):
    """Independent restatement of one simultaneous projection in float64."""
    Dm = (rest[:3] - rest[3]).T.astype(np.float64)  # columns = rest edges
    Ds = (X[:3] - X[3]).T
    Dmi = np.linalg.inv(Dm)
    F = Ds @ Dmi
    I1 = np.sum(F * F)
    dI1 = 2.0 * F @ Dmi.T  # dI1/dDs
    if energy == 4:
        U0, dU0 = I1, dI1
    else:
        IM = I1 - 3.0
        U0 = C[0] * IM + C[1] * IM ** 2 + C[2] * IM ** 3
        dU0 = (C[0] + 2 * C[1] * IM + 3 * C[2] * IM ** 2) * dI1
    J = np.linalg.det(F)
    dJ = J * np.linalg.inv(F).T @ Dmi.T
    U1 = (J - a) ** 2
    dU1 = 2.0 * (J - a) * dJ

    def node_grads(dD):
        g = np.zeros((4, 3))
        g[:3] = dD.T
        g[3] = -g[:3].sum(0)
        return g

    g0, g1 = node_grads(dU0), node_grads(dU1)
    al0, al1 = comp0 / DT ** 2, comp1 / DT ** 2
    A = np.array([[np.sum(w[:, None] * g0 * g0) + 2 * U0 * al0, np.sum(w[:, None] * g0 * g1)],
                  [np.sum(w[:, None] * g0 * g1), np.sum(w[:, None] * g1 * g1) + 2 * U1 * al1]])
    lam = np.linalg.solve(A, [-2 * U0, -2 * U1])
    return X + w[:, None] * (lam[0] * g0 + lam[1] * g1)


def oracle_solve(X, rest, w, energy, comp0, comp1, a, sim=True, variant="f64"):
    el = po.init_elements(rest.reshape(-1), np.arange(4, dtype=np.uint32), 1000.0)
    p = po.make_params(DT, energy=energy, simultaneous=sim)
    Xo = X.reshape(-1).astype(np.float64).copy()
    getattr(po.lib(), "xo_solve_tet_" + variant)(p, Xo, w.astype(np.float32), el["qi"], el["qq"], el["qr"], comp0,
                                                 comp1, a)
    return Xo.reshape(4, 3)


@pytest.mark.parametrize("energy", [4, 7])
@pytest.mark.parametrize("seed", range(6))
def test_projection_matches_textbook(energy, seed):
    rng = np.random.default_rng(seed)
    rest = rest_tet(rng)
    F = np.eye(3) + rng.normal(0, 0.08, (3, 3))
    if energy == 7:
        F = F * 1.05  # keep I1 > 3 so the 1e-4 clamp stays inactive
    X = rest.astype(np.float64) @ F.T + rng.normal(0, 0.01, 3)
    w = rng.uniform(0.5, 2.0, 4)
    comp0, comp1, a = 0.02, 0.003, 1.0
    got = oracle_solve(X, rest, w, energy, comp0, comp1, a)
    want = textbook(X, rest, w, energy, comp0, comp1, a)
    dx = want - X
    assert np.abs(got - want).max() <= 1e-4 * np.abs(dx).max() + 1e-12


def test_textbook_catches_a_planted_mistake():
    """A wrong sign in the volume term is detected by the same comparison."""
    rng = np.random.default_rng(3)
    rest = rest_tet(rng)
    X = rest.astype(np.float64) @ (np.eye(3) + rng.normal(0, 0.08, (3, 3))).T
    w = rng.uniform(0.5, 2.0, 4)
    got = oracle_solve(X, rest, w, 4, 0.02, 0.003, 1.0)
    wrong = textbook(X, rest, w, 4, 0.02, 0.003, -1.0)  # (J + 1)^2 instead of (J - 1)^2
    assert np.abs(got - wrong).max() > 1e-3 * np.abs(wrong - X).max()


@pytest.mark.parametrize("energy", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("sim", [True, False])
def test_projection_conserves_momentum(energy, sim):
    rng = np.random.default_rng(11 + energy)
    rest = rest_tet(rng)
    X = rest.astype(np.float64) @ (1.1 * np.eye(3) + rng.normal(0, 0.05, (3, 3))).T + 0.3
    w = rng.uniform(0.5, 2.0, 4).astype(np.float32).astype(np.float64)  # the oracle takes FP32 w
    m = 1.0 / w
    Xo = oracle_solve(X, rest, w, energy, 0.02, 0.0, 1.0, sim=sim)
    d = Xo - X
    scale = np.abs(d).max() * m.max()
    assert scale > 0
    assert np.abs((m[:, None] * d).sum(0)).max() <= 1e-9 * scale
    if sim:  # one projection: isotropic energies give torque-free corrections
        torque = np.cross(X - X.mean(0), m[:, None] * d).sum(0)
        assert np.abs(torque).max() <= 1e-9 * scale * np.abs(X).max()


@pytest.mark.parametrize("pair", [(3, 4), (5, 7)])
def test_quadrature_energies_equal_prefactored(pair):
    rng = np.random.default_rng(5)
    rest = rest_tet(rng)
    X = rest.astype(np.float64) @ (1.08 * np.eye(3) + rng.normal(0, 0.05, (3, 3))).T
    w = rng.uniform(0.5, 2.0, 4)
    A = oracle_solve(X, rest, w, pair[0], 0.02, 0.0, 1.0)
    B = oracle_solve(X, rest, w, pair[1], 0.02, 0.0, 1.0)
    assert np.abs(A - B).max() <= 1e-7 * np.abs(B - X).max()


def test_mesh_counts_volumes_and_mass():
    W, H, D, h = 7, 3, 2, 0.05
    xyz, tets = xf.tet_block(W, H, D, (h, h, h), (0.0, 0.0, 0.0))
    n, t = xyz.size // 3, tets.size // 4
    assert n == (W + 1) * (H + 1) * (D + 1) and t == 6 * W * H * D
    el = po.init_elements(xyz, tets, 1200.0)
    assert np.allclose(el["vol"], h ** 3 / 6, rtol=1e-5)
    assert np.isclose((1.0 / el["w"]).astype(np.float64).sum(), 1200.0 * W * H * D * h ** 3, rtol=1e-5)
    qi = el["qi"].reshape(-1, 3, 3)  # [tet][column][row]
    P = (xyz.reshape(-1, 3)[tets.reshape(-1, 4)[:, :3]] - xyz.reshape(-1, 3)[tets.reshape(-1, 4)[:, 3:]])
    # Qi is the inverse of the map J whose row n is P_n
    for k in range(0, t, 17):
        Jm = P[k]  # rows = P_n
        Qi = qi[k].T  # row/col matrix from column-major storage
        assert np.allclose(Qi @ Jm, np.eye(3), atol=1e-5)


def test_ground_collision_cases():
    p = po.make_params(DT, collide=1, ground_y=-0.5, friction=0.5)
    o = np.array([0.0, -0.49, 0.0])
    # static: tangential motion 0.001 <= mu * depth = 0.5 * 0.01
    x = np.array([0.001, -0.51, 0.0])
    po.lib().xo_collide_f64(p, 0.0, o, x)
    assert x[1] == -0.5 and x[0] == o[0] and x[2] == o[2]
    # kinetic: tangential 0.1 > 0.005, reduced by exactly mu * depth along its direction
    x = np.array([0.06, -0.51, 0.08])
    po.lib().xo_collide_f64(p, 0.0, o, x)
    assert x[1] == -0.5
    assert np.isclose(np.hypot(x[0], x[2]), 0.1 - 0.005, rtol=1e-12)
    assert np.isclose(x[0] / x[2], 0.06 / 0.08, rtol=1e-12)
    # above the plane: untouched
    x = np.array([0.3, -0.2, 0.1])
    po.lib().xo_collide_f64(p, 0.0, o, x.copy())
    y = x.copy()
    po.lib().xo_collide_f64(p, 0.0, o, y)
    assert np.array_equal(x, y)


def test_sphere_collision_cases():
    p = po.make_params(DT, collide=2, friction=0.5, sphere_c0=(0.0, 0.0, 0.0), sphere_v=(1.0, 0.0, 0.0),
                       sphere_r=0.3)
    t = 0.25
    c = np.array([0.25, 0.0, 0.0])
    rng = np.random.default_rng(2)
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        x = c + d * rng.uniform(0.2, 0.299)
        o = x + rng.normal(0, 0.01, 3)
        po.lib().xo_collide_f64(p, t, o, x)
        assert np.linalg.norm(x - c) >= 0.3 - 1e-6
    # static friction: tangential motion relative to the sphere surface is removed
    x = c + np.array([0.0, 0.295, 0.0])
    n = (x - c) / np.linalg.norm(x - c)  # contact normal
    o = x - np.array([1.0 * DT - 1e-5, 0.0, 0.0])  # moved with the sphere, slightly less
    po.lib().xo_collide_f64(p, t, o, x)
    rel = (x - o) - np.array([1.0, 0.0, 0.0]) * DT
    assert np.abs(rel - n * rel.dot(n)).max() <= 1e-9


def sphere_contact_error(c, r, mu, sv, dt, o, x_in, x_out):
    """How far x_out is from the Coulomb contact of x_in with a moving sphere,
    stated by its geometry (0 when right): the node ends on the tangent plane
    at the closest surface point c + r n; the tangential slip of the substep
    relative to the sphere surface, s = tangential part of (x - o) - sv dt,
    is cancelled when |s| <= mu * depth (static) and otherwise shrinks by
    exactly mu * depth along its own direction (kinetic)."""
    d = x_in - c
    n = d / np.linalg.norm(d)
    depth = r - np.linalg.norm(d)
    surf = c + r * n

    def slip(x):
        rel = (x - o) - sv * dt
        return rel - n * rel.dot(n)

    s_in, s_out = slip(surf), slip(x_out)
    err = abs((x_out - c).dot(n) - r)  # normal coordinate: on the tangent plane
    lim = mu * depth
    if np.linalg.norm(s_in) <= lim:
        err = max(err, np.linalg.norm(s_out))
    else:
        want = s_in * (1.0 - lim / np.linalg.norm(s_in))
        err = max(err, np.linalg.norm(s_out - want))
    return err


def sphere_mutant(kind, c, r, mu, sv, dt, o, x):
    """Plausible mistakes in a sphere contact pass (numpy, for the checks
    above to reject): the sphere's own motion with the wrong sign, the normal
    part left in the friction correction, the depth taken from the squared
    distance, static friction that pins the node to o (right only for a
    sphere at rest)."""
    d = x - c
    dist = np.linalg.norm(d)
    n = d / dist
    depth = (r * r - dist * dist) if kind == "depth_squared" else (r - dist)
    xp = x + n * depth
    sgn = -1.0 if kind == "sv_sign" else 1.0
    rel = (xp - o) - sgn * sv * dt
    rt = rel if kind == "normal_kept" else rel - n * rel.dot(n)
    lim = mu * (r - dist)
    t = np.linalg.norm(rt)
    if t <= lim:
        return o.copy() if kind == "static_to_o" else xp - rt
    return xp - rt * (lim / t)


def _sphere_case(rng, c, r, sv, dt, kinetic):
    """A node inside the sphere near its surface with a previous position o
    whose slip relative to the surface is well above (kinetic) or below
    (static) mu * depth for mu = 0.5."""
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    depth = rng.uniform(2e-5, 2e-4)
    x = c + n * (r - depth)
    tang = rng.normal(size=3)
    tang -= n * tang.dot(n)
    tang /= np.linalg.norm(tang)
    slip = 0.5 * depth * (rng.uniform(1.5, 10.0) if kinetic else rng.uniform(0.05, 0.8))
    # o such that the slip of the pushed-out position relative to the sphere
    # is `slip` along tang, with a normal part of 50-100 um either way
    o = (c + n * r) - sv * dt - tang * slip + n * rng.choice([-1.0, 1.0]) * rng.uniform(5e-5, 1e-4)
    return o, x


@pytest.mark.parametrize("kinetic", [True, False])
def test_sphere_friction_closed_form(kinetic):
    """The sphere branch of the contact pass against the Coulomb geometry,
    at the 96x24x24 beam's regime (0.3 m sphere at 1 m/s, dt = 1/1800, so
    the sphere moves 0.56 mm per substep while contacts are a fraction of
    that deep), over many contact normals; FP64 variant to 1e-12 and the
    float element-math variant (the one the GPU reproduces) to float
    rounding."""
    dt = float(np.float32(1.0) / np.float32(1800.0))
    # parameters the oracle holds as float, so both sides see the same values
    r, mu, sv = float(np.float32(0.3)), 0.5, np.array([1.0, 0.0, 0.0])
    t = 0.125
    c0 = np.array([0.25, -0.125, 0.0625])
    p = po.make_params(dt, collide=2, friction=mu, sphere_c0=tuple(c0), sphere_v=tuple(sv), sphere_r=r)
    cf = c0 + sv * t  # exact in float and double
    rng = np.random.default_rng(40 + kinetic)
    for _ in range(50):
        o, x_in = _sphere_case(rng, cf, r, sv, dt, kinetic)
        x64 = x_in.copy()
        po.lib().xo_collide_f64(p, t, o, x64)
        assert sphere_contact_error(cf, r, mu, sv, dt, o, x_in, x64) <= 1e-12
        x32 = x_in.copy()
        po.lib().xo_collide_ref64(p, t, o, x32)
        # float32 REAL: centre, normal and slip carry ~1e-7 relative rounding of 0.3 m
        assert sphere_contact_error(cf, r, mu, sv, dt, o, x_in, x32) <= 2e-7
        # the check is sharp enough: each planted mistake fails it by far more
        for kind in ("sv_sign", "normal_kept", "depth_squared") + (() if kinetic else ("static_to_o",)):
            bad = sphere_mutant(kind, cf, r, mu, sv, dt, o, x_in)
            assert sphere_contact_error(cf, r, mu, sv, dt, o, x_in, bad) > 1e-6, kind
        # and the numpy restatement without a mistake passes it
        good = sphere_mutant("none", cf, r, mu, sv, dt, o, x_in)
        assert sphere_contact_error(cf, r, mu, sv, dt, o, x_in, good) <= 1e-12


@pytest.mark.parametrize("order", [xf.COLOUR_NATURAL, xf.COLOUR_MINSTD])
@pytest.mark.parametrize("dims,mirrored", [((12, 4, 4), False), ((5, 6, 7), True)])
def test_colouring_valid_and_product_equals_oracle(order, dims, mirrored):
    xyz, tets = xf.tet_block(*dims, (0.1, 0.1, 0.1), (0, 0, 0), mirrored)
    n, t = xyz.size // 3, tets.size // 4
    sweep, colour, nc = xf.mesh_colour(n, tets, order)
    visit = np.arange(t, dtype=np.uint32) if order == xf.COLOUR_NATURAL else po.minstd_order(t)
    onc, ocol = po.colour_greedy(n, tets, visit)
    assert nc == onc and np.array_equal(colour, ocol)
    oorder = np.concatenate([visit[ocol[visit] == c] for c in range(onc)])
    assert np.array_equal(sweep, oorder)
    assert sorted(sweep.tolist()) == list(range(t))
    T4 = tets.reshape(-1, 4)
    valence = np.bincount(T4.reshape(-1), minlength=n).max()
    assert nc >= valence
    for c in range(nc):
        nodes = T4[colour == c].reshape(-1)
        assert np.unique(nodes).size == nodes.size, f"colour {c} shares a node"


@pytest.mark.parametrize("dims", [(12, 4, 4), (96, 24, 24), (5, 7, 3), (3, 3, 3)])
def test_block_24_colouring_valid(dims):
    xyz, tets = xf.tet_block(*dims, (0.1, 0.1, 0.1), (0, 0, 0))
    colour = xf.tet_block_colours(*dims)
    assert colour.max() == 23
    T4 = tets.reshape(-1, 4)
    for c in range(24):
        nodes = T4[colour == c].reshape(-1)
        assert np.unique(nodes).size == nodes.size, f"colour {c} shares a node"
    cnt = np.bincount(colour)
    assert cnt.max() - cnt.min() <= max(1, cnt.max() // 10)


@pytest.mark.parametrize("seed", range(6))
def test_volume_pass_is_pbd_step_of_j_minus_1(seed):
    """SolveVolumeOnly with zero compliance: one projection of U = (J-1)^2 is
    the PBD step of C = J - 1 (lambda = -2U / sum w |dU|^2 gives
    dx = -C w dC / sum w |dC|^2), so J returns to 1 up to second order; with
    compliance the step shrinks by sum w|dU|^2 / (sum w|dU|^2 + 2 U alpha)."""
    rng = np.random.default_rng(100 + seed)
    rest = rest_tet(rng)
    el = po.init_elements(rest.reshape(-1), np.arange(4, dtype=np.uint32), 1000.0)
    w = rng.uniform(0.5, 2.0, 4).astype(np.float32)
    X = rest.astype(np.float64) + rng.normal(0, 0.004, (4, 3))
    Dmi = np.linalg.inv((rest[:3] - rest[3]).T.astype(np.float64))

    def J(Y):
        return np.linalg.det((Y[:3] - Y[3]).T @ Dmi)

    p = po.make_params(DT)

    def residual(scale):
        Xs = rest.astype(np.float64) + scale * (X - rest)
        Ys = Xs.reshape(-1).copy()
        po.lib().xo_solve_volume_f64(p, Ys, w, el["qi"], 0.0)
        return abs(J(Xs) - 1.0), abs(J(Ys.reshape(4, 3)) - 1.0)

    # second order: halving the deformation quarters the remaining volume error
    # (scales keep the error far above the ~1e-8 left by the float32 Qi)
    e0, e1 = residual(0.4)
    h0, h1 = residual(0.2)
    assert e1 < 0.05 * e0
    assert 0.15 < h1 / e1 < 0.35, (e1, h1)
    Y = X.reshape(-1).copy()
    po.lib().xo_solve_volume_f64(p, Y, w, el["qi"], 0.0)
    e0 = abs(J(X) - 1.0)
    # planted mistake: the PBD step with the wrong sign doubles the error
    Z = X.reshape(-1).copy()
    po.lib().xo_solve_volume_f64(p, Z, w, el["qi"], 0.0)
    Zbad = X + (X - Z.reshape(4, 3))
    assert abs(J(Zbad) - 1.0) > 1.5 * e0
    # compliance: the move shrinks by wgg / (wgg + 2 U alpha), from the textbook gradient
    comp = 3e-9
    Yc = X.reshape(-1).copy()
    po.lib().xo_solve_volume_f64(p, Yc, w, el["qi"], comp)
    F = (X[:3] - X[3]).T @ Dmi
    dJ = J(X) * np.linalg.inv(F).T @ Dmi.T
    g = np.zeros((4, 3))
    g[:3] = (2.0 * (J(X) - 1.0) * dJ).T
    g[3] = -g[:3].sum(0)
    wgg = np.sum(w[:, None] * g * g)
    U = (J(X) - 1.0) ** 2
    shrink = wgg / (wgg + 2.0 * U * comp / DT ** 2)
    move0 = np.linalg.norm(Y - X.reshape(-1))
    movec = np.linalg.norm(Yc - X.reshape(-1))
    assert np.isclose(movec / move0, shrink, rtol=1e-9)


def test_ldlt_usage_example():
    A = np.array([1, 2, 3, 4, 5, 6], np.float32)
    b = np.array([7, 8, 9], np.float32)
    xo = np.zeros(3, np.float32)
    po.lib().xo_ldlt_solve(3, A, b, xo)
    assert np.allclose(xo, [-3, 3, 1], atol=1e-5)
    assert np.array_equal(xf.ldlt_solve(A, b), xo)
    # against numpy on a random SPD system
    rng = np.random.default_rng(0)
    M = rng.normal(size=(6, 6))
    S = (M @ M.T + 6 * np.eye(6)).astype(np.float32)
    packed = np.concatenate([S[i, :i + 1] for i in range(6)])
    rhs = rng.normal(size=6).astype(np.float32)
    assert np.allclose(xf.ldlt_solve(packed, rhs), np.linalg.solve(S.astype(np.float64), rhs), rtol=1e-4)
