"""ctypes binding of the C ABI in include/bf/bf.h (libbfamd.so).

The library is the product; this module only marshals numpy arrays across the
ABI for tests, the benchmark and Python callers.  Loading fails loudly when
the HIP library has not been built — there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbfamd.so")

VELOCITY_SETS = {"D2Q9": 2, "D3Q15": 3, "D3Q19": 4, "D3Q27": 5}
Q_OF = {2: 9, 3: 15, 4: 19, 5: 27}
STORAGE = {"fp16": 0, "fp32": 1, "bf16": 2}
BOUNDARY_MODES = {"halfway": 0, "fullway_compat": 1}
PARAMS = {
    "relaxation_time": 0, "constant_force": 1, "interaction_strength": 2,
    "thermal_expansion": 3, "thermal_relaxation_time": 4, "reference_temperature": 5,
    "thermal_cs2": 6, "velocity_limit": 7, "velocity_limit_extreme": 8,
    "shan_chen_synchronous": 9, "tick_pair_planes": 10,
    "tick_sweep_ticks": 11,
}
FIELDS = {"velocity_density": 0, "boundary": 1, "force_temperature": 2,
          "populations_raw": 3, "thermal_raw": 4, "psi": 5, "populations": 6}
STEP_SAVE_MACRO = 1
STEP_CONSERVATION = 2
STATUS = {0: "BF_OK", 1: "BF_ERR_INVALID", 2: "BF_ERR_STATE", 3: "BF_ERR_UNSUPPORTED",
          4: "BF_ERR_DEVICE", 5: "BF_ERR_OOM"}


class BFError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("resolution", C.c_int32 * 3), ("relaxation_time", C.c_float),
                ("periodic", C.c_int32 * 3), ("velocity_set", C.c_int32),
                ("storage", C.c_int32), ("multiphase", C.c_int32),
                ("boundary_mode", C.c_int32), ("device", C.c_int32)]


class Info(C.Structure):
    _fields_ = [("resolution", C.c_int32 * 3), ("q", C.c_int32), ("q_thermal", C.c_int32),
                ("storage", C.c_int32), ("forcing", C.c_int32), ("thermal", C.c_int32),
                ("multiphase", C.c_int32), ("has_boundaries", C.c_int32),
                ("cells", C.c_int64), ("stride", C.c_int64), ("device_bytes", C.c_int64)]


class ChannelConfig(C.Structure):
    _fields_ = [("resolution", C.c_int32 * 3), ("velocity_set", C.c_int32), ("relaxation_time", C.c_float),
                ("inlet_velocity", C.c_float), ("outlet_density", C.c_float), ("device", C.c_int32)]


class AdjointConfig(C.Structure):
    _fields_ = [("steps", C.c_int64), ("checkpoint_every", C.c_int32), ("trajectory_budget_bytes", C.c_int64)]


_lib = None


def _share_hip_runtime_with_torch():
    """libbfamd.so and torch both need libamdhip64.so.7; the first copy loaded
    serves the whole process.  Loading torch first (when it is installed) keeps
    one HIP runtime, so torch.cuda.synchronize() also covers the solver's
    stream and torch keeps seeing the GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library(path=None):
    """Load libbfamd.so (or the library at `path`) and declare every entry
    point's signature.  The library loaded last is the one every Solver and
    module function of this process calls."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    _share_hip_runtime_with_torch()
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build it with boltzmannfluids_amd/build.py (hipcc, gfx950)")
    lib = C.CDLL(p)
    vp, i32, i64, u32, f32p = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.POINTER(C.c_float)
    sig = {
        "bf_abi_version": (i32, []),
        "bf_velocity_count": (i32, [i32]),
        "bf_velocity_table": (i32, [i32, i32, f32p, i32]),
        "bf_create": (i32, [C.POINTER(Config), C.POINTER(vp)]),
        "bf_destroy": (None, [vp]),
        "bf_last_error": (C.c_char_p, [vp]),
        "bf_set_object": (i32, [vp, u32, f32p, f32p, f32p, C.c_float, C.c_float]),
        "bf_clear_objects": (i32, [vp]),
        "bf_init_fields": (i32, [vp, f32p, f32p, f32p, f32p, C.POINTER(C.c_uint8)]),
        "bf_set_param": (i32, [vp, i32, f32p, i32]),
        "bf_get_param": (i32, [vp, i32, f32p, i32]),
        "bf_step": (i32, [vp, i64, u32]),
        "bf_step_timed": (i32, [vp, i64, u32, C.POINTER(C.c_double)]),
        "bf_ticks": (i64, [vp]),
        "bf_synchronize": (i32, [vp]),
        "bf_read_field": (i32, [vp, i32, vp, C.c_size_t]),
        "bf_field_bytes": (C.c_size_t, [vp, i32]),
        "bf_copy_field_to_device": (i32, [vp, i32, vp, C.c_size_t]),
        "bf_conservation": (i32, [vp, C.POINTER(C.c_double)]),
        "bf_last_step_sums": (i32, [vp, C.POINTER(C.c_double)]),
        "bf_get_info": (i32, [vp, C.POINTER(Info)]),
        "bf_bits_per_boundary": (i32, [i32]),
        "bf_pack_boundaries": (i32, [C.POINTER(C.c_uint8), i64, i32, C.POINTER(C.c_uint32)]),
        "bf_unpack_boundaries": (i32, [C.POINTER(C.c_uint32), i64, i32, C.POINTER(C.c_uint8)]),
        "bf_save_mesoscopic": (i32, [vp, C.c_char_p]),
        "bf_save_macroscopic": (i32, [vp, C.c_char_p]),
        "bf_load_mesoscopic": (i32, [vp, C.c_char_p, i64]),
        "bf_load_macroscopic": (i32, [vp, C.c_char_p, i64]),
        "bf_voxelize_mesh": (i32, [i32, f32p, i64, C.POINTER(C.c_uint32), i64, f32p, C.POINTER(C.c_int32),
                                   C.c_uint8, C.POINTER(C.c_uint8), C.POINTER(C.c_int64)]),
        "bf_voxelize_last_error": (C.c_char_p, []),
        "bf_marching_cubes_capacity": (i32, [C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "bf_marching_cubes": (i32, [i32, f32p, C.POINTER(C.c_int32), C.c_float, f32p, f32p, i64,
                                    C.POINTER(C.c_uint32), i64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "bf_isosurface": (i32, [vp, C.c_float, f32p, f32p, i64, C.POINTER(C.c_uint32), i64,
                                C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "bf_marching_cubes_table": (i32, [C.POINTER(C.c_uint8), C.POINTER(C.c_int8)]),
        "bf_marching_cubes_last_error": (C.c_char_p, []),
        "bf_marching_cubes_last_device_ms": (C.c_double, []),
        "bf_channel_create": (i32, [C.POINTER(ChannelConfig), C.POINTER(vp)]),
        "bf_channel_destroy": (None, [vp]),
        "bf_channel_last_error": (C.c_char_p, [vp]),
        "bf_channel_set_porosity": (i32, [vp, f32p]),
        "bf_channel_set_inlet_velocity": (i32, [vp, C.c_float]),
        "bf_channel_forward": (i32, [vp, i64, C.POINTER(C.c_double)]),
        "bf_channel_gradient": (i32, [vp, C.POINTER(AdjointConfig), C.POINTER(C.c_double), f32p,
                                      C.POINTER(C.c_double)]),
        "bf_channel_read_moments": (i32, [vp, f32p, C.c_size_t]),
        "bf_channel_timing": (i32, [vp, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def velocity_table(name_or_id, thermal=False):
    lib = load_library()
    vid = VELOCITY_SETS.get(name_or_id, name_or_id) if not thermal else {"D2Q5": 0, "D3Q7": 1}.get(name_or_id, name_or_id)
    q = {0: 5, 1: 7}[vid] if thermal else Q_OF[vid]
    out = np.zeros(4 * q, np.float32)
    st = lib.bf_velocity_table(vid, 1 if thermal else 0, _fptr(out), out.size)
    if st:
        raise BFError(st, "bf_velocity_table")
    return out.reshape(q, 4)


def bits_per_boundary(object_count):
    return load_library().bf_bits_per_boundary(int(object_count))


def pack_boundaries(ids, bits):
    lib = load_library()
    ids = np.ascontiguousarray(ids, dtype=np.uint8).ravel()
    words = np.zeros((ids.size * bits + 31) // 32, np.uint32)
    st = lib.bf_pack_boundaries(ids.ctypes.data_as(C.POINTER(C.c_uint8)), ids.size, bits,
                                words.ctypes.data_as(C.POINTER(C.c_uint32)))
    if st:
        raise BFError(st, "bf_pack_boundaries")
    return words


def unpack_boundaries(words, n, bits):
    lib = load_library()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    ids = np.zeros(n, np.uint8)
    st = lib.bf_unpack_boundaries(words.ctypes.data_as(C.POINTER(C.c_uint32)), n, bits,
                                  ids.ctypes.data_as(C.POINTER(C.c_uint8)))
    if st:
        raise BFError(st, "bf_unpack_boundaries")
    return ids


def marching_cubes_table():
    """The case table: (triangle_count[256] uint8, edges[256, 15] int8)."""
    lib = load_library()
    n = np.zeros(256, np.uint8)
    e = np.zeros((256, 15), np.int8)
    st = lib.bf_marching_cubes_table(n.ctypes.data_as(C.POINTER(C.c_uint8)), e.ctypes.data_as(C.POINTER(C.c_int8)))
    if st:
        raise BFError(st, "bf_marching_cubes_table")
    return n, e


def _mc_run(call):
    """Size query, then extraction into arrays of the returned sizes."""
    lib = load_library()
    nv, ni = C.c_int64(), C.c_int64()
    st = call(None, 0, None, 0, C.byref(nv), C.byref(ni))
    if st:
        raise BFError(st, lib.bf_marching_cubes_last_error().decode())
    verts = np.zeros((max(nv.value, 1), 3), np.float32)
    idx = np.zeros((max(ni.value, 3) // 3, 3), np.uint32)
    st = call(_fptr(verts), nv.value, idx.ctypes.data_as(C.POINTER(C.c_uint32)), ni.value, C.byref(nv), C.byref(ni))
    if st:
        raise BFError(st, lib.bf_marching_cubes_last_error().decode())
    return verts[: nv.value], idx[: ni.value // 3]


def marching_cubes_device_ms():
    """Device time of the last marching-cubes extraction on this thread (ms)."""
    return load_library().bf_marching_cubes_last_device_ms()


def marching_cubes(volume, threshold, mesh_scale=(1.0, 1.0, 1.0), device=0):
    """Iso-surface of a scalar field given as an array of shape (Z, Y, X) (or
    flat with `volume.shape` set by the caller): returns vertices (V, 3) float32
    and triangles (T, 3) uint32."""
    lib = load_library()
    v = np.ascontiguousarray(volume, np.float32)
    if v.ndim != 3:
        raise ValueError("volume must have shape (Z, Y, X)")
    res = (C.c_int32 * 3)(v.shape[2], v.shape[1], v.shape[0])
    sc = np.ascontiguousarray(mesh_scale, np.float32)
    return _mc_run(lambda vp_, vc, ip, ic, nv, ni: lib.bf_marching_cubes(
        int(device), _fptr(v), res, float(threshold), _fptr(sc), vp_, vc, ip, ic, nv, ni))


def voxelize_mesh(vertices, triangles, resolution, value=1, model=None, volume=None, device=0):
    """Mark the voxels inside a closed mesh (vertices (V, 3), triangles (T, 3))
    with `value` in a uint8 volume of shape (Z, Y, X) (a new zero volume when
    None); `model` is a 4x4 matrix applied to the vertices (row-major numpy
    layout, i.e. model @ [x, y, z, 1]).  Returns the volume."""
    lib = load_library()
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    X, Y, Z = (int(r) for r in resolution)
    if volume is None:
        volume = np.zeros((Z, Y, X), np.uint8)
    if volume.shape != (Z, Y, X) or volume.dtype != np.uint8 or not volume.flags.c_contiguous:
        raise ValueError("volume must be a contiguous uint8 array of shape (Z, Y, X)")
    m = np.eye(4, dtype=np.float32) if model is None else np.asarray(model, np.float32)
    mc_ = np.ascontiguousarray(m.T)      # column-major for the C ABI
    res = (C.c_int32 * 3)(X, Y, Z)
    n = C.c_int64()
    st = lib.bf_voxelize_mesh(int(device), _fptr(v), len(v), t.ctypes.data_as(C.POINTER(C.c_uint32)), len(t),
                              _fptr(mc_), res, int(value), volume.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(n))
    if st:
        raise BFError(st, lib.bf_voxelize_last_error().decode())
    return volume


class Solver:
    """One solver handle (device lattice + stream) behind the C ABI."""

    def __init__(self, resolution, relaxation_time, periodic=(True, True, True),
                 velocity_set="D3Q19", storage="fp32", multiphase=False,
                 boundary_mode="halfway", device=0):
        self._lib = load_library()
        cfg = Config()
        cfg.resolution[:] = [int(r) for r in resolution]
        cfg.relaxation_time = float(relaxation_time)
        cfg.periodic[:] = [1 if p else 0 for p in periodic]
        cfg.velocity_set = VELOCITY_SETS.get(velocity_set, velocity_set)
        cfg.storage = STORAGE.get(storage, storage)
        cfg.multiphase = 1 if multiphase else 0
        cfg.boundary_mode = BOUNDARY_MODES.get(boundary_mode, boundary_mode)
        cfg.device = int(device)
        h = C.c_void_p()
        st = self._lib.bf_create(C.byref(cfg), C.byref(h))
        if st:
            raise BFError(st, self._lib.bf_last_error(None).decode())
        self._h = h
        self.resolution = tuple(int(r) for r in resolution)
        self.q = Q_OF[cfg.velocity_set]
        self.cells = int(np.prod(self.resolution))

    def _check(self, st, what):
        if st:
            raise BFError(st, f"{what}: {self._lib.bf_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_object(self, oid, velocity=(0, 0, 0), angular_velocity=(0, 0, 0),
                   center_of_mass=(0, 0, 0), temperature=1.0, effective_density=1.0):
        v = _f32(velocity, 3)
        a = _f32(angular_velocity, 3)
        c = _f32(center_of_mass, 3)
        self._check(self._lib.bf_set_object(self._h, int(oid), _fptr(v), _fptr(a), _fptr(c),
                                            float(temperature), float(effective_density)), "bf_set_object")

    def clear_objects(self):
        self._check(self._lib.bf_clear_objects(self._h), "bf_clear_objects")

    def init_fields(self, rho, u, force=None, temperature=None, boundary_id=None):
        n = self.cells
        rho = _f32(rho, n)
        u = _f32(u, 3 * n)
        force = _f32(force, 3 * n)
        temperature = _f32(temperature, n)
        ids = None
        if boundary_id is not None:
            ids = np.ascontiguousarray(boundary_id, dtype=np.uint8).ravel()
            if ids.size != n:
                raise ValueError("boundary_id size mismatch")
        self._check(self._lib.bf_init_fields(
            self._h, _fptr(rho), _fptr(u), _fptr(force), _fptr(temperature),
            None if ids is None else ids.ctypes.data_as(C.POINTER(C.c_uint8))), "bf_init_fields")

    def set_param(self, name, value):
        v = _f32(np.atleast_1d(value))
        self._check(self._lib.bf_set_param(self._h, PARAMS.get(name, name), _fptr(v), v.size), "bf_set_param")

    def get_param(self, name, n=1):
        v = np.zeros(max(n, 1), np.float32)
        self._check(self._lib.bf_get_param(self._h, PARAMS.get(name, name), _fptr(v), v.size), "bf_get_param")
        return v if n > 1 else float(v[0])

    def step(self, n=1, save=False, conservation=False):
        flags = (STEP_SAVE_MACRO if save else 0) | (STEP_CONSERVATION if conservation else 0)
        self._check(self._lib.bf_step(self._h, int(n), flags), "bf_step")

    def step_timed(self, n, save=False, conservation=False):
        flags = (STEP_SAVE_MACRO if save else 0) | (STEP_CONSERVATION if conservation else 0)
        ms = C.c_double()
        self._check(self._lib.bf_step_timed(self._h, int(n), flags, C.byref(ms)), "bf_step_timed")
        return ms.value

    def synchronize(self):
        self._check(self._lib.bf_synchronize(self._h), "bf_synchronize")

    @property
    def ticks(self):
        return int(self._lib.bf_ticks(self._h))

    def info(self):
        i = Info()
        self._check(self._lib.bf_get_info(self._h, C.byref(i)), "bf_get_info")
        return {k: (list(getattr(i, k)) if k == "resolution" else getattr(i, k)) for k, _ in Info._fields_}

    # checkpoint files: <directory>/tick_<ticks>.bfmeso / .bfmacro
    def save_mesoscopic(self, directory):
        self._check(self._lib.bf_save_mesoscopic(self._h, os.fsencode(directory)), "bf_save_mesoscopic")

    def save_macroscopic(self, directory):
        self._check(self._lib.bf_save_macroscopic(self._h, os.fsencode(directory)), "bf_save_macroscopic")

    def load_mesoscopic(self, directory, tick):
        self._check(self._lib.bf_load_mesoscopic(self._h, os.fsencode(directory), int(tick)), "bf_load_mesoscopic")

    def load_macroscopic(self, directory, tick):
        self._check(self._lib.bf_load_macroscopic(self._h, os.fsencode(directory), int(tick)), "bf_load_macroscopic")

    def isosurface(self, threshold, mesh_scale=(1.0, 1.0, 1.0)):
        """Marching-cubes iso-surface of the density saved on the last save
        tick, extracted on the device: vertices (V, 3) and triangles (T, 3)."""
        sc = np.ascontiguousarray(mesh_scale, np.float32)
        return _mc_run(lambda vp_, vc, ip, ic, nv, ni: self._lib.bf_isosurface(
            self._h, float(threshold), _fptr(sc), vp_, vc, ip, ic, nv, ni))

    def copy_field_to_device(self, field, address, nbytes):
        """Copies a field in the encoding of the reference's texture for it
        (velocity_density / force_temperature: float4 per cell; boundary: one
        byte per cell, 0 fluid and 255 object) into device memory at `address`,
        device to device, on the solver's stream."""
        fid = FIELDS.get(field, field)
        self._check(self._lib.bf_copy_field_to_device(self._h, fid, C.c_void_p(address), nbytes),
                    "bf_copy_field_to_device")

    def read(self, field):
        fid = FIELDS.get(field, field)
        nbytes = self._lib.bf_field_bytes(self._h, fid)
        if nbytes == 0:
            raise BFError(1, f"field {field} does not exist for this solver")
        dtype = np.uint8 if fid == 1 else np.float32
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype)
        self._check(self._lib.bf_read_field(self._h, fid, out.ctypes.data_as(C.c_void_p), nbytes), "bf_read_field")
        n = self.cells
        if fid in (0, 2):
            return out.reshape(n, 4)
        if fid in (3, 4, 6):
            return out.reshape(-1, n)
        return out

    def conservation(self):
        out = np.zeros(5, np.float64)
        self._check(self._lib.bf_conservation(self._h, out.ctypes.data_as(C.POINTER(C.c_double))), "bf_conservation")
        return out

    def last_step_sums(self):
        out = np.zeros(5, np.float64)
        self._check(self._lib.bf_last_step_sums(self._h, out.ctypes.data_as(C.POINTER(C.c_double))), "bf_last_step_sums")
        return out


class Channel:
    """Differentiable channel (bf_channel_*): forward runs and adjoint gradients."""

    def __init__(self, resolution, relaxation_time, inlet_velocity, velocity_set="D2Q9", outlet_density=1.0, device=0):
        self._lib = load_library()
        cfg = ChannelConfig()
        cfg.resolution[:] = [int(r) for r in resolution]
        cfg.velocity_set = VELOCITY_SETS.get(velocity_set, velocity_set)
        cfg.relaxation_time = float(relaxation_time)
        cfg.inlet_velocity = float(inlet_velocity)
        cfg.outlet_density = float(outlet_density)
        cfg.device = int(device)
        h = C.c_void_p()
        st = self._lib.bf_channel_create(C.byref(cfg), C.byref(h))
        if st:
            raise BFError(st, self._lib.bf_channel_last_error(None).decode())
        self._h = h
        self.resolution = tuple(int(r) for r in resolution)
        self.cells = int(np.prod(self.resolution))

    def _check(self, st, what):
        if st:
            raise BFError(st, f"{what}: {self._lib.bf_channel_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bf_channel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_porosity(self, gamma):
        g = _f32(gamma, self.cells)
        self._check(self._lib.bf_channel_set_porosity(self._h, _fptr(g)), "bf_channel_set_porosity")

    def set_inlet_velocity(self, u):
        self._check(self._lib.bf_channel_set_inlet_velocity(self._h, float(u)), "bf_channel_set_inlet_velocity")

    def forward(self, steps):
        J = C.c_double()
        self._check(self._lib.bf_channel_forward(self._h, int(steps), C.byref(J)), "bf_channel_forward")
        return J.value

    def gradient(self, steps, checkpoint_every=100, trajectory_budget_bytes=0):
        a = AdjointConfig()
        a.steps = int(steps)
        a.checkpoint_every = int(checkpoint_every)
        a.trajectory_budget_bytes = int(trajectory_budget_bytes)
        J = C.c_double()
        du = C.c_double()
        dg = np.zeros(self.cells, np.float32)
        self._check(self._lib.bf_channel_gradient(self._h, C.byref(a), C.byref(J), _fptr(dg), C.byref(du)),
                    "bf_channel_gradient")
        return J.value, dg, du.value

    def moments(self):
        out = np.zeros(4 * self.cells, np.float32)
        self._check(self._lib.bf_channel_read_moments(self._h, _fptr(out), out.nbytes), "bf_channel_read_moments")
        return out.reshape(self.cells, 4)

    def timing(self):
        out = np.zeros(3, np.float64)
        self._check(self._lib.bf_channel_timing(self._h, out.ctypes.data_as(C.POINTER(C.c_double))), "bf_channel_timing")
        return out
