"""Python binding of libxpbdfem.so (the C ABI declared in include/xpbdfem.h).

Thin ctypes layer used by the tests and the benchmark. All computation runs
in the HIP library; this module only marshals numpy arrays. There is no CPU
fallback: loading fails loudly when the library is missing, and creating a
context fails when no GPU is visible.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxpbdfem.so")

ENERGY_MIXED, ENERGY_MIXED_SEL, ENERGY_YEOH_SKIN, ENERGY_YEOH_SKIN_SEL, ENERGY_YEOH_SKIN_FAST = 3, 4, 5, 6, 7
# vertex-ring ("continuous") energies: the sweep runs over vertex rings
ENERGY_CONTINUOUS_PIXAR, ENERGY_CONTINUOUS_MIXED, ENERGY_CONTINUOUS_SKIN = 8, 9, 10
ENERGY_PIXAR, ENERGY_PIXAR_REDUCED, ENERGY_PIXAR_SEL = 0, 1, 2
ENERGY_CUBE_NEO, ENERGY_CUBE_SKIN = 11, 12  # no constraint on tets
COLLIDE_GROUND, COLLIDE_SPHERE = 1, 2
RAYLEIGH_PAPER, RAYLEIGH_LIMIT, RAYLEIGH_POST, RAYLEIGH_POST_AMORTIZED = 0, 1, 2, 3
COLOUR_NATURAL, COLOUR_MINSTD = 0, 1
LAUNCH_PER_COLOUR, LAUNCH_PERSISTENT, LAUNCH_BLOCKS = 0, 1, 2
SKIN_YEOH = (0.1095, 14.95, 4.595)  # Fem.cpp:526

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class XfError(RuntimeError):
    pass


class Settings(C.Structure):
    _fields_ = [
        ("energy", C.c_uint32),
        ("simultaneous", C.c_uint32),
        ("compliance", C.c_float),
        ("poissons_ratio", C.c_float),
        ("gravity", C.c_float * 2),
        ("drag", C.c_float),
        ("damping", C.c_float),
        ("pbd_damping", C.c_float),
        ("volume_passes", C.c_uint32),
        ("yeoh", C.c_float * 3),
        ("collide", C.c_uint32),
        ("ground_y", C.c_float),
        ("friction", C.c_float),
        ("sphere_c0", C.c_float * 3),
        ("sphere_v", C.c_float * 3),
        ("sphere_r", C.c_float),
        ("rayleigh_type", C.c_uint32),
        ("pbd_damping_amortized", C.c_float),
        ("unlock_pins", C.c_uint32),
    ]


class BodyDesc(C.Structure):
    _fields_ = [
        ("xyz", C.c_void_p),
        ("n_nodes", C.c_uint32),
        ("tets", C.c_void_p),
        ("n_tets", C.c_uint32),
        ("density", C.c_float),
        ("pinned", C.c_void_p),
        ("velocity", C.c_void_p),
        ("tet_mu", C.c_void_p),
        ("tet_colour", C.c_void_p),
        ("xyz64", C.c_void_p),
        ("pin_first_substep", C.c_uint32),
    ]


class SceneOpts(C.Structure):
    _fields_ = [("colour_order", C.c_uint32), ("block_size", C.c_uint32), ("use_graph", C.c_uint32),
                ("launch_mode", C.c_uint32), ("persistent_blocks", C.c_uint32), ("xcd_slabs", C.c_uint32),
                ("ring_lanes", C.c_uint32), ("sweep_blocks", C.c_uint32), ("block_lanes", C.c_uint32),
                ("separate_finish", C.c_uint32), ("tet_lanes", C.c_uint32), ("contract", C.c_uint32)]


class SceneInfo(C.Structure):
    _fields_ = [("n_bodies", C.c_uint32), ("n_nodes", C.c_uint32), ("n_tets", C.c_uint32),
                ("n_colours", C.c_uint32), ("device_bytes", C.c_uint64), ("n_rings", C.c_uint32),
                ("n_ring_colours", C.c_uint32), ("n_blocks", C.c_uint32), ("block_launches", C.c_uint32),
                ("n_block_nodes", C.c_uint32), ("block_lanes", C.c_uint32), ("n_block_slots", C.c_uint32),
                ("fused_finish", C.c_uint32), ("tet_lanes", C.c_uint32),
                ("block_lds_limit", C.c_uint32), ("contract", C.c_uint32), ("block_node_stride", C.c_uint32)]


_lib = None


def loaded_hip_runtimes():
    """Distinct libamdhip64 files mapped into this process.

    PyTorch wheels ship their own HIP runtime. It is shared with this library
    only when torch is imported before libxpbdfem.so loads; the other order
    maps two runtimes, which must not both drive the device."""
    found = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                path = line.split()[-1]
                if "libamdhip64" in os.path.basename(path):
                    found.add(os.path.realpath(path))
    except OSError:
        pass
    return sorted(found)


def lib():
    """Loads libxpbdfem.so from the package directory (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise XfError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "xf_version": ([], C.c_char_p),
        "xf_settings_default": ([C.POINTER(Settings)], None),
        "xf_ctx_create": ([C.c_int, C.POINTER(vp)], C.c_int),
        "xf_ctx_destroy": ([vp], None),
        "xf_last_error": ([vp], C.c_char_p),
        "xf_scene_create": ([vp, C.POINTER(vp)], C.c_int),
        "xf_scene_destroy": ([vp], None),
        "xf_scene_add_body": ([vp, C.POINTER(BodyDesc), C.POINTER(C.c_uint32)], C.c_int),
        "xf_scene_finalize": ([vp, C.POINTER(SceneOpts)], C.c_int),
        "xf_scene_info_get": ([vp, C.POINTER(SceneInfo)], C.c_int),
        "xf_scene_set_settings": ([vp, C.c_uint32, C.POINTER(Settings)], C.c_int),
        "xf_scene_set_tet_mu": ([vp, C.c_uint32, _f32p], C.c_int),
        "xf_scene_step": ([vp, C.c_uint32, C.c_float], C.c_int),
        "xf_scene_sync": ([vp], C.c_int),
        "xf_scene_get_state": ([vp, C.c_uint32, vp, vp], C.c_int),
        "xf_scene_set_state": ([vp, C.c_uint32, vp, vp], C.c_int),
        "xf_scene_get_time": ([vp, C.POINTER(C.c_float)], C.c_int),
        "xf_scene_set_time": ([vp, C.c_float], C.c_int),
        "xf_scene_get_tick": ([vp, C.POINTER(C.c_uint32)], C.c_int),
        "xf_mesh_read_tetgen": ([C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 vp, vp], C.c_int),
        "xf_mesh_tets_from_stream": ([vp, C.c_uint32, C.POINTER(C.c_uint32), vp], C.c_int),
        "xf_mesh_error": ([], C.c_char_p),
        "xf_scene_set_tick": ([vp, C.c_uint32], C.c_int),
        "xf_scene_get_order": ([vp, C.c_uint32, vp, vp], C.c_int),
        "xf_scene_get_ring_order": ([vp, C.c_uint32, vp, vp, vp], C.c_int),
        "xf_mesh_rings": ([C.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp, vp], C.c_int),
        "xf_mesh_sweep_blocks": ([C.c_uint32, vp, C.c_uint32, vp, C.c_uint32, vp, vp, vp, vp, vp], C.c_int),
        "xf_scene_get_elements": ([vp, C.c_uint32, vp, vp, vp, vp, vp], C.c_int),
        "xf_scene_profile": ([vp, C.c_uint32, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float),
                              C.POINTER(C.c_uint32)], C.c_int),
        "xf_scene_count_nonfinite": ([vp, C.POINTER(C.c_uint32)], C.c_int),
        "xf_scene_time_steps": ([vp, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(C.c_float)], C.c_int),
        "xf_scene_set_kinematic": ([vp, C.c_uint32, C.c_uint32, vp, vp], C.c_int),
        "xf_scene_set_manip": ([vp, C.c_uint32, C.c_uint32, _f32p], C.c_int),
        "xf_scene_adjoint": ([vp, C.c_uint32, C.c_float, C.c_uint32, _u32p, vp, C.c_uint32, C.POINTER(C.c_double),
                              vp, vp], C.c_int),
        "xf_scene_adjoint_noalloc": ([vp, C.c_uint32, C.c_float, C.c_uint32, _u32p, vp, C.c_uint32,
                                      C.POINTER(C.c_double), vp, vp], C.c_int),
        "xf_scene_adjoint_fp64": ([vp, C.c_uint32, C.c_float, C.c_uint32, _u32p, vp, C.c_uint32,
                                   C.POINTER(C.c_double), vp, vp], C.c_int),
        "xf_scene_autotune": ([vp, C.c_uint32, C.c_float, C.c_uint32, C.POINTER(SceneOpts), C.POINTER(C.c_float)],
                              C.c_int),
        "xf_mesh_tet_block": ([C.c_uint32, C.c_uint32, C.c_uint32, _f32p, _f32p, C.c_uint32, _f32p, _u32p],
                              C.c_int),
        "xf_minstd_order": ([C.c_uint32, _u32p], C.c_int),
        "xf_mesh_tet_block_colours": ([C.c_uint32, C.c_uint32, C.c_uint32, _u32p], C.c_int),
        "xf_normal_fill": ([C.c_uint64, C.c_float, C.c_uint32, _f32p], C.c_int),
        "xf_ldlt_solve": ([C.c_uint32, _f32p, _f32p, _f32p], C.c_int),
        "xf_mesh_colour": ([C.c_uint32, _u32p, C.c_uint32, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_uint32)],
                           C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def exported_symbols():
    """Names the binding expects the library to export."""
    return [
        "xf_version", "xf_settings_default", "xf_ctx_create", "xf_ctx_destroy", "xf_last_error", "xf_scene_create",
        "xf_scene_destroy", "xf_scene_add_body", "xf_scene_finalize", "xf_scene_info_get", "xf_scene_set_settings",
        "xf_scene_set_tet_mu", "xf_scene_step", "xf_scene_sync", "xf_scene_get_state", "xf_scene_set_state",
        "xf_scene_get_time", "xf_scene_set_time", "xf_scene_get_order", "xf_scene_get_ring_order",
        "xf_scene_get_elements", "xf_scene_profile", "xf_scene_count_nonfinite", "xf_mesh_tet_block", "xf_minstd_order", "xf_normal_fill", "xf_ldlt_solve",
        "xf_mesh_colour", "xf_scene_autotune", "xf_scene_adjoint", "xf_mesh_tet_block_colours",
        "xf_scene_get_tick", "xf_scene_set_tick", "xf_mesh_read_tetgen", "xf_mesh_tets_from_stream", "xf_mesh_error",
        "xf_mesh_rings", "xf_mesh_sweep_blocks", "xf_scene_time_steps", "xf_scene_adjoint_noalloc",
        "xf_scene_set_kinematic", "xf_scene_set_manip", "xf_scene_adjoint_fp64",
    ]


# -------------------------------------------------------------- utilities

def tet_block(W, H, D, scale, origin, mirrored=False):
    n = (W + 1) * (H + 1) * (D + 1)
    xyz = np.zeros(3 * n, np.float32)
    tets = np.zeros(24 * W * H * D, np.uint32)
    rc = lib().xf_mesh_tet_block(W, H, D, np.asarray(scale, np.float32), np.asarray(origin, np.float32),
                                 int(mirrored), xyz, tets)
    if rc:
        raise XfError(f"xf_mesh_tet_block failed ({rc})")
    return xyz, tets


def tet_block_colours(W, H, D):
    """The 24-colouring of a uniform-pattern tet block."""
    out = np.zeros(6 * W * H * D, np.uint32)
    if lib().xf_mesh_tet_block_colours(W, H, D, out):
        raise XfError("xf_mesh_tet_block_colours failed")
    return out


def mesh_rings(tets):
    """Vertex rings of a tet mesh (the reference's OneRings) as a dict of CSR
    arrays: node_off, nodes, tet_off, tets, idx (4 per ring tet), largest."""
    tets = np.ascontiguousarray(tets, np.uint32)
    t = tets.size // 4
    nr, tot, big = C.c_uint32(), C.c_uint32(), C.c_uint32()
    rc = lib().xf_mesh_rings(t, tets.ctypes.data, C.byref(nr), C.byref(tot), None, None, None, None, None,
                             C.byref(big))
    if rc:
        raise XfError(f"xf_mesh_rings failed ({rc})")
    out = dict(node_off=np.zeros(nr.value + 1, np.uint32), nodes=np.zeros(tot.value, np.uint32),
               tet_off=np.zeros(nr.value + 1, np.uint32), tets=np.zeros(4 * t, np.uint32),
               idx=np.zeros(16 * t, np.uint8))
    lib().xf_mesh_rings(t, tets.ctypes.data, C.byref(nr), C.byref(tot), out["node_off"].ctypes.data,
                        out["nodes"].ctypes.data, out["tet_off"].ctypes.data, out["tets"].ctypes.data,
                        out["idx"].ctypes.data, C.byref(big))
    out["largest"] = big.value
    return out


def mesh_sweep_blocks(xyz, tets, colour_order=COLOUR_NATURAL):
    """Sweep blocks of a tet mesh as the scene builds them with sweep_blocks:
    dict of block (per tet), phase (per tet, inside its block), launch (per
    block, 0..3) and max_phases."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    tets = np.ascontiguousarray(tets, np.uint32)
    n, t = xyz.size // 3, tets.size // 4
    nb, mp = C.c_uint32(), C.c_uint32()
    block, phase, launch = np.zeros(t, np.uint32), np.zeros(t, np.uint32), np.zeros(max(t, 1), np.uint32)
    rc = lib().xf_mesh_sweep_blocks(n, xyz.ctypes.data, t, tets.ctypes.data, colour_order, C.byref(nb),
                                    block.ctypes.data, phase.ctypes.data, launch.ctypes.data, C.byref(mp))
    if rc:
        raise XfError(f"xf_mesh_sweep_blocks failed ({rc}): {lib().xf_mesh_error().decode()}")
    return dict(block=block, phase=phase, launch=launch[:nb.value].copy(), max_phases=mp.value)


def autoresize(xyz):
    """GeoLinear3d::Init's autoResize (Geo.cpp:712-729): FP64 positions
    0.085 (x - centre) / largest bounding-box extent, with the box taken in
    float and the arithmetic in double, as the reference computes them."""
    p = np.asarray(xyz, np.float32).reshape(-1, 3)
    mn, mx = p.min(axis=0), p.max(axis=0)
    dims = mx - mn
    max_dim = float(max(dims[0], max(dims[1], dims[2])))
    centre = 0.5 * (mn + mx).astype(np.float64)
    return (0.085 * ((p.astype(np.float64) - centre) / max_dim)).reshape(-1)


def minstd_order(n):
    o = np.zeros(n, np.uint32)
    lib().xf_minstd_order(n, o)
    return o


def normal_fill(seed, sigma, n):
    out = np.zeros(n, np.float32)
    lib().xf_normal_fill(seed, sigma, n, out)
    return out


def mesh_colour(n_nodes, tets, colour_order=COLOUR_NATURAL):
    """Host-side greedy colouring: (sweep order, colour per tet, colour count)."""
    tets = np.ascontiguousarray(tets, np.uint32)
    t = tets.size // 4
    order = np.zeros(t, np.uint32)
    colour = np.zeros(t, np.uint32)
    nc = C.c_uint32()
    rc = lib().xf_mesh_colour(n_nodes, tets, t, colour_order, order, colour, C.byref(nc))
    if rc:
        raise XfError(f"xf_mesh_colour failed ({rc})")
    return order, colour, nc.value


def read_tetgen(node_path, ele_path, orient=True):
    """(xyz float32[3n], tets uint32[4t]) from a tetgen .node / .ele pair."""
    n, t = C.c_uint32(), C.c_uint32()
    L = lib()
    rc = L.xf_mesh_read_tetgen(str(node_path).encode(), str(ele_path).encode(), int(orient), C.byref(n), C.byref(t),
                               None, None)
    if rc:
        raise XfError(f"xf_mesh_read_tetgen failed ({rc}): {L.xf_mesh_error().decode()}")
    xyz = np.zeros(3 * n.value, np.float32)
    tets = np.zeros(4 * t.value, np.uint32)
    rc = L.xf_mesh_read_tetgen(str(node_path).encode(), str(ele_path).encode(), int(orient), C.byref(n), C.byref(t),
                               xyz.ctypes.data, tets.ctypes.data)
    if rc:
        raise XfError(f"xf_mesh_read_tetgen failed ({rc}): {L.xf_mesh_error().decode()}")
    return xyz, tets


def write_tetgen(prefix, xyz, tets, base=0):
    """Writes prefix.node / prefix.ele in tetgen's format (numbering from base)."""
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    tets = np.asarray(tets, np.int64).reshape(-1, 4)
    with open(f"{prefix}.node", "w") as f:
        f.write(f"{len(xyz)} 3 0 0\n")
        for i, p in enumerate(xyz):
            f.write(f"{i + base} {float(p[0]):.17g} {float(p[1]):.17g} {float(p[2]):.17g}\n")
    with open(f"{prefix}.ele", "w") as f:
        f.write(f"{len(tets)} 4 0\n")
        for k, t in enumerate(tets):
            f.write(f"{k + base} {t[0] + base} {t[1] + base} {t[2] + base} {t[3] + base}\n")
    return f"{prefix}.node", f"{prefix}.ele"


def tets_from_stream(idx):
    """The [4, i0, i1, i2, i3]* index stream of GeoLinear3d::Init to tets."""
    idx = np.ascontiguousarray(idx, np.uint32)
    t = C.c_uint32()
    L = lib()
    rc = L.xf_mesh_tets_from_stream(idx.ctypes.data, idx.size, C.byref(t), None)
    if rc:
        raise XfError(f"xf_mesh_tets_from_stream failed ({rc}): {L.xf_mesh_error().decode()}")
    tets = np.zeros(4 * t.value, np.uint32)
    rc = L.xf_mesh_tets_from_stream(idx.ctypes.data, idx.size, C.byref(t), tets.ctypes.data)
    if rc:
        raise XfError(f"xf_mesh_tets_from_stream failed ({rc}): {L.xf_mesh_error().decode()}")
    return tets


def ldlt_solve(A_packed, b):
    A = np.ascontiguousarray(A_packed, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    x = np.zeros_like(b)
    rc = lib().xf_ldlt_solve(b.size, A, b, x)
    if rc:
        raise XfError("xf_ldlt_solve failed")
    return x


def default_settings(**kw):
    s = Settings()
    lib().xf_settings_default(C.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        fld = getattr(s, k)
        if isinstance(fld, C.Array):
            fld[:] = v
        else:
            setattr(s, k, v)
    return s


# ---------------------------------------------------------------- objects

class Context:
    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.xf_ctx_create(device, C.byref(h))
        if rc:
            raise XfError(f"xf_ctx_create(device={device}) failed ({rc}): no usable HIP device")
        self.h = h
        self.scenes = weakref.WeakSet()

    def check(self, rc, what):
        if rc:
            msg = lib().xf_last_error(self.h).decode()
            raise XfError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self.h:
            for sc in list(self.scenes):
                sc.close()
            lib().xf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().xf_scene_create(ctx.h, C.byref(h)), "xf_scene_create")
        self.h = h
        ctx.scenes.add(self)
        self.bodies = []  # (n_nodes, n_tets)
        self._keep = []

    def add_body(self, xyz, tets, density, pinned=None, velocity=None, tet_mu=None, tet_colour=None, xyz64=None,
                 pin_first_substep=False):
        """xyz: float32 positions; xyz64 (optional): FP64 positions that replace
        them (e.g. autoresized ones, see autoresize()); pin_first_substep: pins
        keep inverse mass 1/m until their first locked substep."""
        tets = np.ascontiguousarray(tets, np.uint32).reshape(-1)
        d = BodyDesc()
        d.pin_first_substep = int(bool(pin_first_substep))
        if xyz64 is not None:
            xyz = np.ascontiguousarray(xyz64, np.float64).reshape(-1)
            d.xyz64 = xyz.ctypes.data
        else:
            xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
            d.xyz = xyz.ctypes.data
        d.n_nodes = xyz.size // 3
        d.tets = tets.ctypes.data
        d.n_tets = tets.size // 4
        d.density = density
        keep = [xyz, tets]
        if pinned is not None:
            p = np.ascontiguousarray(pinned, np.uint8).reshape(-1)
            d.pinned = p.ctypes.data
            keep.append(p)
        if velocity is not None:
            v = np.ascontiguousarray(velocity, np.float32).reshape(-1)
            d.velocity = v.ctypes.data
            keep.append(v)
        if tet_mu is not None:
            m = np.ascontiguousarray(tet_mu, np.float32).reshape(-1)
            d.tet_mu = m.ctypes.data
            keep.append(m)
        if tet_colour is not None:
            c = np.ascontiguousarray(tet_colour, np.uint32).reshape(-1)
            d.tet_colour = c.ctypes.data
            keep.append(c)
        bid = C.c_uint32()
        self.ctx.check(lib().xf_scene_add_body(self.h, C.byref(d), C.byref(bid)), "xf_scene_add_body")
        self.bodies.append((d.n_nodes, d.n_tets))
        return bid.value

    def finalize(self, colour_order=COLOUR_NATURAL, block_size=64, use_graph=True, launch_mode=LAUNCH_PER_COLOUR,
                 persistent_blocks=0, xcd_slabs=False, ring_lanes=0, sweep_blocks=False, block_lanes=0,
                 separate_finish=False, tet_lanes=0, contract=False):
        o = SceneOpts()
        o.tet_lanes = tet_lanes
        o.contract = int(bool(contract))
        o.sweep_blocks = int(sweep_blocks)
        o.block_lanes = block_lanes
        o.separate_finish = int(bool(separate_finish))
        o.xcd_slabs = int(xcd_slabs)
        o.ring_lanes = ring_lanes
        o.colour_order = colour_order
        o.block_size = block_size
        o.use_graph = 1 if use_graph else 0
        o.launch_mode = launch_mode
        o.persistent_blocks = persistent_blocks
        self.ctx.check(lib().xf_scene_finalize(self.h, C.byref(o)), "xf_scene_finalize")

    def info(self):
        i = SceneInfo()
        self.ctx.check(lib().xf_scene_info_get(self.h, C.byref(i)), "xf_scene_info_get")
        return dict(n_bodies=i.n_bodies, n_nodes=i.n_nodes, n_tets=i.n_tets, n_colours=i.n_colours,
                    device_bytes=i.device_bytes, n_rings=i.n_rings, n_ring_colours=i.n_ring_colours,
                    n_blocks=i.n_blocks, block_launches=i.block_launches, n_block_nodes=i.n_block_nodes,
                    block_lanes=i.block_lanes, n_block_slots=i.n_block_slots,
                    fused_finish=i.fused_finish, tet_lanes=i.tet_lanes, block_lds_limit=i.block_lds_limit,
                    contract=i.contract, block_node_stride=i.block_node_stride)

    def set_settings(self, body, settings=None, **kw):
        s = settings if settings is not None else default_settings(**kw)
        self.ctx.check(lib().xf_scene_set_settings(self.h, body, C.byref(s)), "xf_scene_set_settings")

    def set_tet_mu(self, body, mu):
        mu = np.ascontiguousarray(mu, np.float32)
        self.ctx.check(lib().xf_scene_set_tet_mu(self.h, body, mu), "xf_scene_set_tet_mu")

    def step(self, substeps, dt):
        self.ctx.check(lib().xf_scene_step(self.h, substeps, dt), "xf_scene_step")

    def sync(self):
        self.ctx.check(lib().xf_scene_sync(self.h), "xf_scene_sync")

    def get_state(self, body=0):
        n = self.bodies[body][0]
        x = np.zeros(3 * n, np.float64)
        v = np.zeros(3 * n, np.float64)
        self.ctx.check(lib().xf_scene_get_state(self.h, body, x.ctypes.data, v.ctypes.data), "xf_scene_get_state")
        return x, v

    def set_state(self, body, x=None, v=None):
        xp = np.ascontiguousarray(x, np.float64) if x is not None else None
        vp = np.ascontiguousarray(v, np.float64) if v is not None else None
        self.ctx.check(lib().xf_scene_set_state(self.h, body, xp.ctypes.data if xp is not None else None,
                                                vp.ctypes.data if vp is not None else None), "xf_scene_set_state")

    def time(self):
        t = C.c_float()
        self.ctx.check(lib().xf_scene_get_time(self.h, C.byref(t)), "xf_scene_get_time")
        return t.value

    def set_time(self, t):
        self.ctx.check(lib().xf_scene_set_time(self.h, t), "xf_scene_set_time")

    def tick(self):
        k = C.c_uint32()
        self.ctx.check(lib().xf_scene_get_tick(self.h, C.byref(k)), "xf_scene_get_tick")
        return k.value

    def set_tick(self, k):
        self.ctx.check(lib().xf_scene_set_tick(self.h, k), "xf_scene_set_tick")

    def order(self, body=0):
        t = self.bodies[body][1]
        o = np.zeros(t, np.uint32)
        c = np.zeros(t, np.uint32)
        self.ctx.check(lib().xf_scene_get_order(self.h, body, o.ctypes.data, c.ctypes.data), "xf_scene_get_order")
        return o, c

    def ring_order(self, body=0):
        """(vertex rings of the body in sweep order, colour of each ring)."""
        n = C.c_uint32()
        self.ctx.check(lib().xf_scene_get_ring_order(self.h, body, C.byref(n), None, None), "xf_scene_get_ring_order")
        o = np.zeros(n.value, np.uint32)
        c = np.zeros(n.value, np.uint32)
        self.ctx.check(lib().xf_scene_get_ring_order(self.h, body, C.byref(n), o.ctypes.data, c.ctypes.data),
                       "xf_scene_get_ring_order")
        return o, c

    def elements(self, body=0):
        n, t = self.bodies[body]
        out = dict(qi=np.zeros(9 * t, np.float32), qq=np.zeros(3 * t, np.float32), qr=np.zeros(3 * t, np.float32),
                   vol=np.zeros(t, np.float32), w=np.zeros(n, np.float32))
        self.ctx.check(lib().xf_scene_get_elements(self.h, body, out["qi"].ctypes.data, out["qq"].ctypes.data,
                                                   out["qr"].ctypes.data, out["vol"].ctypes.data,
                                                   out["w"].ctypes.data), "xf_scene_get_elements")
        return out

    def profile(self, substeps, dt):
        mp, mn, nl = C.c_float(), C.c_float(), C.c_uint32()
        self.ctx.check(lib().xf_scene_profile(self.h, substeps, dt, C.byref(mp), C.byref(mn), C.byref(nl)),
                       "xf_scene_profile")
        return mp.value, mn.value, nl.value

    def set_kinematic(self, body, nodes, xyz):
        """Nodes held at FP64 targets after every sweep (LockRight); empty clears."""
        nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        self.ctx.check(lib().xf_scene_set_kinematic(self.h, body, nodes.size, nodes.ctypes.data, xyz.ctypes.data),
                       "xf_scene_set_kinematic")

    def set_manip(self, body, node, target=(0.0, 0.0, 0.0)):
        """The manipulation constraint on one node (node None: off)."""
        t = np.ascontiguousarray(target, np.float32)
        self.ctx.check(lib().xf_scene_set_manip(self.h, body, 0xFFFFFFFF if node is None else node, t),
                       "xf_scene_set_manip")

    def time_steps(self, calls, substeps, dt):
        """Device ms of `calls` step calls between two HIP events on the
        library stream (synchronous)."""
        ms = C.c_float()
        self.ctx.check(lib().xf_scene_time_steps(self.h, calls, substeps, dt, C.byref(ms)), "xf_scene_time_steps")
        return ms.value

    def autotune(self, substeps, dt, reps=3):
        """Picks the fastest launch shape; returns (launch_mode, size, ms per call),
        size = workgroup size (per-colour, sweep blocks) or grid blocks
        (persistent)."""
        o = SceneOpts()
        ms = C.c_float()
        self.ctx.check(lib().xf_scene_autotune(self.h, substeps, dt, reps, C.byref(o), C.byref(ms)),
                       "xf_scene_autotune")
        size = o.persistent_blocks if o.launch_mode == LAUNCH_PERSISTENT else o.block_size
        return o.launch_mode, size, ms.value

    def adjoint(self, substeps, dt, loss_at, targets, window=20, per_body=False, noalloc=False, fp64=False):
        """Forward rollout + adjoint: returns (loss, dL/dlog mu per tet, all
        bodies concatenated), plus each body's loss term if per_body.
        targets: [n_loss, total nodes * 3] float64. noalloc: through
        xf_scene_adjoint_noalloc (device scratch of fixed size, no
        allocation). fp64: through xf_scene_adjoint_fp64 (the FP64 element
        model)."""
        if noalloc and fp64:
            raise ValueError("the allocation-free adjoint runs the FP32 element model only")
        loss_at = np.ascontiguousarray(loss_at, np.uint32)
        targets = np.ascontiguousarray(targets, np.float64)
        n_tets = sum(t for _, t in self.bodies)
        grad = np.zeros(n_tets, np.float64)
        body_loss = np.zeros(len(self.bodies), np.float64)
        L = C.c_double()
        name = "xf_scene_adjoint_noalloc" if noalloc else ("xf_scene_adjoint_fp64" if fp64 else "xf_scene_adjoint")
        self.ctx.check(getattr(lib(), name)(self.h, substeps, dt, loss_at.size, loss_at, targets.ctypes.data,
                                            window, C.byref(L), grad.ctypes.data, body_loss.ctypes.data), name)
        return (L.value, grad, body_loss) if per_body else (L.value, grad)

    def count_nonfinite(self):
        c = C.c_uint32()
        self.ctx.check(lib().xf_scene_count_nonfinite(self.h, C.byref(c)), "xf_scene_count_nonfinite")
        return c.value

    def close(self):
        if self.h:
            if self.ctx.h:  # xf_ctx_destroy already freed the scene otherwise
                lib().xf_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
