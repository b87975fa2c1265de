"""ctypes wrapper over the oracle libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker. The product never imports
this module.

  liboracle.so  C restatement of the reference substep (xf_oracle.c)
  libxfref.so   the unmodified reference compiled from its own sources plus
                ref_harness.cpp (present only where it was built)
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libxfref.so")
# the same reference sources at -O3 -march=x86-64-v3 (timing only)
REF_FAST_SO = os.path.join(HERE, "_ref", "libxfref_fast.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class XoParams(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("gravity", C.c_float * 2),
        ("drag", C.c_float),
        ("energy", C.c_int),
        ("simultaneous", C.c_int),
        ("yeoh", C.c_float * 3),
        ("collide", C.c_int),
        ("ground_y", C.c_float),
        ("friction", C.c_float),
        ("sphere_c0", C.c_float * 3),
        ("sphere_v", C.c_float * 3),
        ("sphere_r", C.c_float),
        ("volume_passes", C.c_int),
        ("compliance", C.c_float),
        ("damping", C.c_float),
        ("rayleigh_type", C.c_int),
        ("pbd", C.c_float),
        ("pbd_amortized", C.c_float),
        ("tick", C.c_uint32),
        ("poisson", C.c_float),
    ]


class XoRings(C.Structure):
    _fields_ = [
        ("n_rings", C.c_uint32),
        ("node_off", C.c_void_p),
        ("nodes", C.c_void_p),
        ("tet_off", C.c_void_p),
        ("tets", C.c_void_p),
        ("idx", C.c_void_p),
    ]


class XrSettings(C.Structure):
    _fields_ = [
        ("compliance", C.c_float),
        ("poissons_ratio", C.c_float),
        ("gravity", C.c_float * 2),
        ("drag", C.c_float),
        ("energy", C.c_uint32),
        ("simultaneous", C.c_uint32),
        ("volume_passes", C.c_uint32),
        ("damping", C.c_float),
        ("rayleigh_type", C.c_uint32),
        ("pbd_damping", C.c_float),
        ("volume_pbd", C.c_float),
        ("amortized_volume_pbd", C.c_float),
        ("tick_id", C.c_uint32),
    ]


SKIN_YEOH = (0.1095, 14.95, 4.595)  # Fem.cpp:526


def make_params(dt, gravity=(0.0, -9.81), drag=0.0, energy=4, simultaneous=True, yeoh=SKIN_YEOH, collide=0,
                ground_y=0.0, friction=0.0, sphere_c0=(0, 0, 0), sphere_v=(0, 0, 0), sphere_r=0.0, volume_passes=0,
                compliance=0.0, damping=0.0, rayleigh_type=0, pbd=0.0, pbd_amortized=0.0, poisson=0.5):
    p = XoParams()
    p.dt = dt
    p.gravity[:] = gravity
    p.drag = drag
    p.energy = energy
    p.simultaneous = 1 if simultaneous else 0
    p.yeoh[:] = yeoh
    p.collide = collide
    p.ground_y = ground_y
    p.friction = friction
    p.sphere_c0[:] = sphere_c0
    p.sphere_v[:] = sphere_v
    p.sphere_r = sphere_r
    p.volume_passes = volume_passes
    p.compliance = compliance
    p.damping = damping
    p.rayleigh_type = rayleigh_type
    p.pbd = pbd
    p.pbd_amortized = pbd_amortized
    p.poisson = poisson
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        L.xo_tet_block.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _f32p, _f32p, C.c_int, _f32p, _u32p]
        L.xo_minstd_order.argtypes = [C.c_uint32, _u32p]
        L.xo_minstd_ring_order.argtypes = [C.c_uint32, C.c_uint32, _u32p]
        L.xo_init_elements.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32, C.c_float, _f32p, _f32p, _f32p, _f32p,
                                       _f32p]
        L.xo_material.argtypes = [C.c_uint32, _f32p, C.c_float, C.c_void_p, C.c_float, _f32p, _f32p, _f32p]
        L.xo_colour_greedy.argtypes = [C.c_uint32, _u32p, C.c_uint32, _u32p, _u32p]
        L.xo_colour_greedy.restype = C.c_uint32
        for sfx, sp in (("_ref64", _f64p), ("_f32", _f32p), ("_f64", _f64p)):
            fn = getattr(L, "xo_substep" + sfx)
            fn.argtypes = [C.POINTER(XoParams), C.c_uint32, sp, sp, sp, _f32p, C.c_void_p, C.c_uint32, _u32p,
                           _u32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
            fn = getattr(L, "xo_solve_volume" + sfx)
            fn.argtypes = [C.POINTER(XoParams), sp, _f32p, _f32p, C.c_float]
            fn = getattr(L, "xo_pbd_damp" + sfx)
            fn.argtypes = [sp, sp, _f32p, C.c_float]
            fn = getattr(L, "xo_rayleigh_tet" + sfx)
            fn.argtypes = [C.POINTER(XoParams), sp, sp, _f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float,
                           C.c_float]
            fn = getattr(L, "xo_adjoint_tet" + sfx)
            fn.argtypes = [C.POINTER(XoParams), sp, sp, _f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float]
            fn.restype = C.c_double if sfx == "_f64" else C.c_float
            fn = getattr(L, "xo_rollout_adjoint" + sfx)
            fn.argtypes = [C.POINTER(XoParams), C.c_uint32, sp, sp, _f32p, C.c_void_p, C.c_uint32, _u32p, _u32p,
                           _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_uint32, C.c_uint32, _u32p, sp, C.c_uint32,
                           _f64p]
            fn.restype = C.c_double
            fn = getattr(L, "xo_collide" + sfx)
            fn.argtypes = [C.POINTER(XoParams), C.c_float, sp, sp]
            fn = getattr(L, "xo_solve_tet" + sfx)
            fn.argtypes = [C.POINTER(XoParams), sp, _f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float]
            fn = getattr(L, "xo_substep_rings" + sfx)
            fn.argtypes = [C.POINTER(XoParams), C.c_uint32, sp, sp, sp, _f32p, C.c_void_p, C.POINTER(XoRings), _u32p,
                           C.c_uint32, _u32p, _u32p, _f32p, _f32p, _f32p, _f32p, C.c_float]
            fn = getattr(L, "xo_ring_ug" + sfx)
            fn.argtypes = [sp, C.POINTER(XoRings), C.c_uint32, _f32p, _f32p, _f32p, _f64p, _f64p, _f64p]
        L.xo_build_rings.argtypes = [C.c_uint32, _u32p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.xo_build_rings.restype = C.c_uint32
        L.xo_ldlt_solve.argtypes = [C.c_uint32, _f32p, _f32p, _f32p]
        L.xo_surface_area.argtypes = [_f32p, _u32p, C.c_uint32, _f32p]
        _lib = L
    return _lib


_ref = {}


def have_ref(fast=False):
    return os.path.exists(REF_FAST_SO if fast else REF_SO)


def ref(fast=False):
    """The compiled reference with its harness: the parity build, or with
    fast the -O3 -march=x86-64-v3 build (for timing)."""
    if fast not in _ref:
        path = REF_FAST_SO if fast else REF_SO
        if not os.path.exists(path):
            raise RuntimeError(f"reference harness missing: {path}")
        L = C.CDLL(path)
        L.xr_create.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32, C.c_float]
        L.xr_create.restype = C.c_void_p
        L.xr_create_ex.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32, C.c_float, C.c_int]
        L.xr_create_ex.restype = C.c_void_p
        L.xr_destroy.argtypes = [C.c_void_p]
        L.xr_vert_count.argtypes = [C.c_void_p]
        L.xr_vert_count.restype = C.c_uint32
        L.xr_element_count.argtypes = [C.c_void_p]
        L.xr_element_count.restype = C.c_uint32
        L.xr_get_order.argtypes = [C.c_void_p, _u32p]
        L.xr_set_order.argtypes = [C.c_void_p, _u32p]
        L.xr_set_pins.argtypes = [C.c_void_p, _u8p]
        L.xr_get_elements.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, _f32p, _f32p]
        L.xr_get_tets.argtypes = [C.c_void_p, _u32p]
        L.xr_get_area.argtypes = [C.c_void_p, _f32p]
        L.xr_armadillo_t4.argtypes = [C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32),
                                      C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint32)]
        L.xr_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.xr_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.xr_substeps.argtypes = [C.c_void_p, C.POINTER(XrSettings), C.c_float, C.c_uint32]
        L.xr_solve_tet.argtypes = [C.POINTER(XrSettings), C.c_float, _f64p, _f32p, _f32p, _f32p, _f32p, C.c_float]
        L.xr_gen_tet_block.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                       C.c_uint32, C.c_void_p, C.c_void_p]
        L.xr_gen_tet_block.restype = C.c_uint32
        L.xr_ldlt.argtypes = [C.c_uint32, _f32p, _f32p, _f32p]
        L.xr_volume.argtypes = [C.c_void_p]
        L.xr_volume.restype = C.c_float
        L.xr_ring_count.argtypes = [C.c_void_p]
        L.xr_ring_count.restype = C.c_uint32
        L.xr_get_vorder.argtypes = [C.c_void_p, _u32p]
        L.xr_set_vorder.argtypes = [C.c_void_p, _u32p]
        L.xr_get_ring.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.xr_substeps_many.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32]
        _ref[fast] = L
    return _ref[fast]


# ------------------------------------------------------------------ helpers

def tet_block(W, H, D, scale, origin, mirrored=False):
    n = (W + 1) * (H + 1) * (D + 1)
    xyz = np.zeros(3 * n, np.float32)
    tets = np.zeros(24 * W * H * D, np.uint32)
    lib().xo_tet_block(W, H, D, np.asarray(scale, np.float32), np.asarray(origin, np.float32), int(mirrored), xyz,
                       tets)
    return xyz, tets


def minstd_order(n):
    o = np.zeros(n, np.uint32)
    lib().xo_minstd_order(n, o)
    return o


def minstd_ring_order(n_tets, n_rings):
    """The reference's vOrder for a mesh of n_tets tets and n_rings vertices."""
    o = np.zeros(n_rings, np.uint32)
    lib().xo_minstd_ring_order(n_tets, n_rings, o)
    return o


def init_elements(xyz, tets, density):
    n = xyz.size // 3
    t = tets.size // 4
    qi = np.zeros(9 * t, np.float32)
    qq = np.zeros(3 * t, np.float32)
    qr = np.zeros(3 * t, np.float32)
    vol = np.zeros(t, np.float32)
    w = np.zeros(n, np.float32)
    lib().xo_init_elements(xyz, n, tets, t, density, qi, qq, qr, vol, w)
    return dict(qi=qi, qq=qq, qr=qr, vol=vol, w=w)


def material(vol, compliance=None, tet_mu=None, nu=0.5):
    t = vol.size
    c0 = np.zeros(t, np.float32)
    c1 = np.zeros(t, np.float32)
    a = np.zeros(t, np.float32)
    mu_ptr = None
    if tet_mu is not None:
        tet_mu = np.ascontiguousarray(tet_mu, np.float32)
        mu_ptr = tet_mu.ctypes.data
    lib().xo_material(t, vol, 0.0 if compliance is None else compliance, mu_ptr, nu, c0, c1, a)
    return c0, c1, a


def colour_greedy(n_nodes, tets, visit):
    t = tets.size // 4
    col = np.zeros(t, np.uint32)
    nc = lib().xo_colour_greedy(n_nodes, tets, t, np.ascontiguousarray(visit, np.uint32), col)
    return nc, col


class Rings:
    """Vertex rings of a tet mesh (xo_build_rings: the reference's OneRing
    construction) in CSR arrays, plus the XoRings view the oracle takes."""

    def __init__(self, tets):
        tets = np.ascontiguousarray(tets, np.uint32)
        t = tets.size // 4
        nr, tot = C.c_uint32(), C.c_uint32()
        if lib().xo_build_rings(t, tets, C.byref(nr), C.byref(tot), None, None, None, None, None) == 0:
            raise ValueError("a vertex ring has more than 256 nodes")
        self.n_rings = nr.value
        self.node_off = np.zeros(self.n_rings + 1, np.uint32)
        self.nodes = np.zeros(max(tot.value, 1), np.uint32)
        self.tet_off = np.zeros(self.n_rings + 1, np.uint32)
        self.tets = np.zeros(max(4 * t, 1), np.uint32)
        self.idx = np.zeros(max(16 * t, 1), np.uint8)
        self.largest = lib().xo_build_rings(t, tets, C.byref(nr), C.byref(tot), self.node_off.ctypes.data,
                                            self.nodes.ctypes.data, self.tet_off.ctypes.data, self.tets.ctypes.data,
                                            self.idx.ctypes.data)
        self.view = XoRings(self.n_rings, self.node_off.ctypes.data, self.nodes.ctypes.data,
                            self.tet_off.ctypes.data, self.tets.ctypes.data, self.idx.ctypes.data)

    def ring(self, v):
        """(nodes, tets, idx[k, 4]) of ring v."""
        a, b = self.node_off[v], self.node_off[v + 1]
        c, d = self.tet_off[v], self.tet_off[v + 1]
        return self.nodes[a:b].copy(), self.tets[c:d].copy(), self.idx[4 * c:4 * d].reshape(-1, 4).copy()


class OracleBody:
    """Node state + element constants of one body for the oracle substep."""

    def __init__(self, xyz, tets, density, pinned=None, compliance=None, tet_mu=None, nu=0.5, variant="f32",
                 velocity=None):
        self.variant = variant
        st = np.float32 if variant == "f32" else np.float64
        self.n = xyz.size // 3
        self.t = tets.size // 4
        self.tets = np.ascontiguousarray(tets, np.uint32)
        el = init_elements(xyz, self.tets, density)
        self.qi, self.qq, self.qr, self.vol, self.w = el["qi"], el["qq"], el["qr"], el["vol"], el["w"]
        self.pinned = None
        if pinned is not None:
            self.pinned = np.ascontiguousarray(pinned, np.uint8)
            self.w = self.w.copy()
            self.w[self.pinned != 0] = 0.0
        self.c0, self.c1, self.a = material(self.vol, compliance, tet_mu, nu)
        self.X = xyz.astype(st).copy()
        self.O = self.X.copy()
        self.V = np.zeros_like(self.X)
        if velocity is not None:
            self.V[:] = np.asarray(velocity, np.float32).reshape(-1)
        self.clock = np.float32(0.0)
        self.tick = 0  # Settings::tickId of the next substep
        self.area = np.zeros(self.t, np.float32)
        lib().xo_surface_area(np.ascontiguousarray(xyz, np.float32), self.tets, self.t, self.area)

    def substep(self, params, order, count=1, ring_order=None):
        """`count` substeps; the collider clock advances by dt in float32 per
        substep (as the device clock does) and the substep sees its end time.
        Vertex-ring energies (params.energy 8, 9, 10) sweep rings in
        ring_order; `order` is then the tet order of the PBD damping pass."""
        rings = 8 <= params.energy <= 10
        if rings:
            if ring_order is None:
                raise ValueError("vertex-ring energies need a ring order")
            if getattr(self, "rings", None) is None:
                self.rings = Rings(self.tets)
            fn = getattr(lib(), "xo_substep_rings_" + self.variant)
            ring_order = np.ascontiguousarray(ring_order, np.uint32)
        else:
            fn = getattr(lib(), "xo_substep_" + self.variant)
        pin = self.pinned.ctypes.data if self.pinned is not None else None
        order = np.ascontiguousarray(order, np.uint32)
        dt = np.float32(params.dt)
        for _ in range(count):
            self.clock = np.float32(self.clock + dt)
            params.tick = self.tick
            if rings:
                fn(C.byref(params), self.n, self.X, self.O, self.V, self.w, pin, C.byref(self.rings.view),
                   ring_order, self.t, order, self.tets, self.qq, self.qr, self.vol, self.area, float(self.clock))
            else:
                fn(C.byref(params), self.n, self.X, self.O, self.V, self.w, pin, self.t, order, self.tets, self.qi,
                   self.qq, self.qr, self.c0, self.c1, self.a, self.vol, self.area, float(self.clock))
            self.tick += 1


def params_from_settings(dt, settings):
    """xo_params from the keyword settings of a workload."""
    s = dict(settings)
    return make_params(float(dt), gravity=s.get("gravity", (0.0, -9.81)), drag=s.get("drag", 0.0),
                       energy=s.get("energy", 4), simultaneous=bool(s.get("simultaneous", 1)),
                       yeoh=s.get("yeoh", SKIN_YEOH), collide=s.get("collide", 0), ground_y=s.get("ground_y", 0.0),
                       friction=s.get("friction", 0.0), sphere_c0=s.get("sphere_c0", (0, 0, 0)),
                       sphere_v=s.get("sphere_v", (0, 0, 0)), sphere_r=s.get("sphere_r", 0.0),
                       volume_passes=s.get("volume_passes", 0), compliance=s.get("compliance", 0.0) or 0.0,
                       damping=s.get("damping", 0.0), rayleigh_type=s.get("rayleigh_type", 0),
                       pbd=s.get("pbd_damping", 0.0), pbd_amortized=s.get("pbd_damping_amortized", 0.0),
                       poisson=s.get("poissons_ratio", 0.5))


def body_from_workload(wl, variant="f32"):
    """OracleBody with a workload's mesh, pins, velocity and material."""
    s = wl["settings"]
    return OracleBody(wl["xyz"], wl["tets"], wl["density"], wl["pinned"], compliance=s.get("compliance"),
                      tet_mu=wl["tet_mu"], nu=s.get("poissons_ratio", 0.5), variant=variant,
                      velocity=wl["velocity"])


def ref_armadillo_t4():
    """The compiled reference's Armadillo tet mesh: (xyz float32, index
    stream uint32), copied out of the reference build (test use only)."""
    nodes, idx = C.POINTER(C.c_float)(), C.POINTER(C.c_uint32)()
    nn, ni = C.c_uint32(), C.c_uint32()
    ref().xr_armadillo_t4(C.byref(nodes), C.byref(nn), C.byref(idx), C.byref(ni))
    return np.ctypeslib.as_array(nodes, (nn.value,)).copy(), np.ctypeslib.as_array(idx, (ni.value,)).copy()


class RefBody:
    """The compiled reference's GeoLinear3d driven through ref_harness."""

    def __init__(self, xyz, tets, density, pinned=None, order=None, autoresize=False, fast=False):
        self.fast = fast
        L = ref(fast)
        self.n = xyz.size // 3
        self.t = tets.size // 4
        self.h = L.xr_create_ex(np.ascontiguousarray(xyz, np.float32), self.n,
                                np.ascontiguousarray(tets, np.uint32), self.t, density, int(autoresize))
        L.xr_set_pins(self.h, np.ascontiguousarray(pinned if pinned is not None else np.zeros(self.n), np.uint8))
        if order is not None:
            L.xr_set_order(self.h, np.ascontiguousarray(order, np.uint32))

    def __del__(self):
        if getattr(self, "h", None):
            ref(self.fast).xr_destroy(self.h)
            self.h = None

    def settings(self, compliance, nu=0.5, gravity=(0.0, -9.81), drag=0.0, energy=4, simultaneous=True,
                 volume_passes=0, damping=0.0, rayleigh_type=0, pbd_damping=0.0, volume_pbd=0.0,
                 amortized_volume_pbd=0.0, tick_id=0):
        s = XrSettings()
        s.volume_passes = volume_passes
        s.damping = damping
        s.rayleigh_type = rayleigh_type
        s.pbd_damping = pbd_damping
        s.volume_pbd = volume_pbd
        s.amortized_volume_pbd = amortized_volume_pbd
        s.tick_id = tick_id
        s.compliance = compliance
        s.poissons_ratio = nu
        s.gravity[:] = gravity
        s.drag = drag
        s.energy = energy
        s.simultaneous = 1 if simultaneous else 0
        return s

    def substeps(self, settings, dt, count):
        ref(self.fast).xr_substeps(self.h, C.byref(settings), dt, count)

    def state(self):
        X = np.zeros(3 * self.n)
        V = np.zeros(3 * self.n)
        ref(self.fast).xr_get_state(self.h, X.ctypes.data, V.ctypes.data)
        return X, V

    def set_state(self, X=None, V=None):
        Xp = np.ascontiguousarray(X, np.float64) if X is not None else None
        Vp = np.ascontiguousarray(V, np.float64) if V is not None else None
        ref(self.fast).xr_set_state(self.h, Xp.ctypes.data if Xp is not None else None,
                           Vp.ctypes.data if Vp is not None else None)

    def elements(self):
        qi = np.zeros(9 * self.t, np.float32)
        qq = np.zeros(3 * self.t, np.float32)
        qr = np.zeros(3 * self.t, np.float32)
        vol = np.zeros(self.t, np.float32)
        w = np.zeros(self.n, np.float32)
        ref(self.fast).xr_get_elements(self.h, qi, qq, qr, vol, w)
        return dict(qi=qi, qq=qq, qr=qr, vol=vol, w=w)

    def area(self):
        a = np.zeros(self.t, np.float32)
        ref(self.fast).xr_get_area(self.h, a)
        return a

    def order(self):
        o = np.zeros(self.t, np.uint32)
        ref(self.fast).xr_get_order(self.h, o)
        return o

    def ring_order(self):
        """GeoLinear3d::vOrder (one entry per ring)."""
        o = np.zeros(ref(self.fast).xr_ring_count(self.h), np.uint32)
        ref(self.fast).xr_get_vorder(self.h, o)
        return o

    def set_ring_order(self, order):
        ref(self.fast).xr_set_vorder(self.h, np.ascontiguousarray(order, np.uint32))

    def ring(self, v):
        """(nodes, tets, idx[k, 4]) of the reference's OneRing of vertex v."""
        nn, nt = C.c_uint32(), C.c_uint32()
        ref(self.fast).xr_get_ring(self.h, v, C.byref(nn), C.byref(nt), None, None, None)
        nodes = np.zeros(nn.value, np.uint32)
        tets = np.zeros(nt.value, np.uint32)
        idx = np.zeros(4 * nt.value, np.uint16)
        ref(self.fast).xr_get_ring(self.h, v, C.byref(nn), C.byref(nt), nodes.ctypes.data, tets.ctypes.data, idx.ctypes.data)
        return nodes, tets, idx.reshape(-1, 4)


def substeps_many(bodies, settings, dt, count, threads):
    """`count` reference substeps of every RefBody in `bodies` (one settings
    object each), spread over `threads` host threads (all of one build)."""
    fast = bodies[0].fast
    hs = (C.c_void_p * len(bodies))(*[b.h for b in bodies])
    ss = (XrSettings * len(bodies))(*settings)
    ref(fast).xr_substeps_many(hs, ss, len(bodies), dt, count, threads)


def rollout_loss(body, params, order, n_sub, loss_at, targets):
    """Forward-only loss of xo_rollout_adjoint (oracle/xo_substep.inc):
    L = sum_j (1/n) sum_i |X_i(after substep loss_at[j]) - target_j,i|^2,
    each squared-error sum taken serially in node-component order; `body`
    is advanced n_sub substeps. For finite differences of the adjoint."""
    targets = np.asarray(targets).reshape(len(loss_at), -1)
    L, done = 0.0, 0
    for j, la in enumerate(loss_at):
        body.substep(params, order, count=int(la) - done)
        done = int(la)
        d = (body.X.astype(np.float64) - targets[j]).tolist()
        e = 0.0
        for x in d:
            e += x * x
        L += e / float(body.n)
    if n_sub > done:
        body.substep(params, order, count=int(n_sub) - done)
    return L


def rollout_adjoint(body, params, order, n_sub, loss_at, targets, window=20):
    """Forward rollout of `body` (an OracleBody; its X, V are advanced) and the
    adjoint of the mean-squared position loss at substeps `loss_at`.
    Returns (L, dL/dlog(comp0) per tet)."""
    fn = getattr(lib(), "xo_rollout_adjoint_" + body.variant)
    st = body.X.dtype
    loss_at = np.ascontiguousarray(loss_at, np.uint32)
    targets = np.ascontiguousarray(np.asarray(targets).reshape(-1), st)
    glog = np.zeros(body.t, np.float64)
    pin = body.pinned.ctypes.data if body.pinned is not None else None
    L = fn(C.byref(params), body.n, body.X, body.V, body.w, pin, body.t, np.ascontiguousarray(order, np.uint32),
           body.tets, body.qi, body.qq, body.qr, body.c0, body.c1, body.a, n_sub, loss_at.size, loss_at, targets,
           window, glog)
    if L < 0:
        raise ValueError("rollout adjoint rejected its input")
    return L, glog
