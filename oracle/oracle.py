"""ctypes wrapper of the CPU oracle (bf_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the solver package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"f64": os.path.join(HERE, "liboracle_f64.so"), "f32": os.path.join(HERE, "liboracle_f32.so")}
VELOCITY_SETS = {"D2Q9": 2, "D3Q15": 3, "D3Q19": 4, "D3Q27": 5}
STORAGE = {"fp16": 0, "fp32": 1, "bf16": 2}
PARAMS = {
    "relaxation_time": 0, "constant_force": 1, "interaction_strength": 2,
    "thermal_expansion": 3, "thermal_relaxation_time": 4, "reference_temperature": 5,
    "thermal_cs2": 6, "velocity_limit": 7, "velocity_limit_extreme": 8,
    "shan_chen_synchronous": 9,
}
_cache = {}


def build():
    """Compile both oracle libraries with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(precision="f64"):
    if precision in _cache:
        return _cache[precision]
    path = LIBS[precision]
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    vp, dp, i32, i64 = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_long
    sig = {
        "orc_velocity_table": (i32, [i32, i32, dp]),
        "orc_rest_fix_f32": (C.c_float, [i32, i32]),
        "orc_opposite": (i32, [i32]),
        "orc_threads": (i32, []),
        "orc_create": (vp, [i32, i32, i32, i32, i32, i32, i32, i32, i32, i32, C.c_double]),
        "orc_destroy": (None, [vp]),
        "orc_set_param": (i32, [vp, i32, dp, i32]),
        "orc_set_object": (i32, [vp, i32, dp, C.c_double, C.c_double]),
        "orc_init_fields": (i32, [vp, dp, dp, dp, dp, C.POINTER(C.c_uint8)]),
        "orc_step": (i32, [vp, i64, C.c_uint]),
        "orc_conservation": (None, [vp, dp]),
        "orc_last_sums": (None, [vp, dp]),
        "orc_ticks": (i64, [vp]),
        "orc_q": (i32, [vp]),
        "orc_qt": (i32, [vp]),
        "orc_forcing": (i32, [vp]),
        "orc_thermal": (i32, [vp]),
        "orc_read_field": (i32, [vp, i32, dp]),
        "orc_write_populations": (None, [vp, dp]),
        "orc_aa_maps": (None, [vp, i32, C.POINTER(C.c_long), C.POINTER(C.c_long)]),
        "orc_two_lattice": (None, [vp, dp, i64]),
        "orc_bits_per_boundary": (i32, [i32]),
        "orc_pack": (None, [C.POINTER(C.c_uint8), i64, i32, C.POINTER(C.c_uint32)]),
        "orc_unpack": (None, [C.POINTER(C.c_uint32), i64, i32, C.POINTER(C.c_uint8)]),
        "orc_half_rne": (C.c_float, [C.c_float]),
        "orc_bf16_rne": (C.c_float, [C.c_float]),
        "och_create": (vp, [i32, i32, i32, i32, C.c_double, C.c_double, C.c_double]),
        "och_destroy": (None, [vp]),
        "och_set_porosity": (None, [vp, dp]),
        "och_set_inlet": (None, [vp, C.c_double]),
        "och_profile": (None, [vp, dp]),
        "och_forward": (C.c_double, [vp, i64]),
        "och_moments": (None, [vp, dp]),
        "och_precollision": (None, [vp, dp, i32, i32, i32, dp]),
        "och_gradient": (C.c_double, [vp, i64, dp, dp]),
        "och_tick_pair": (None, [vp, dp, dp, dp, dp]),
        "och_gradient_checkpointed": (C.c_double, [vp, i64, i64, i32, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _cache[precision] = lib
    return lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def velocity_table(name, thermal=False, precision="f64"):
    lib = load(precision)
    vid = {"D2Q5": 0, "D3Q7": 1}[name] if thermal else VELOCITY_SETS[name]
    out = np.zeros(4 * 27, np.float64)
    q = lib.orc_velocity_table(vid, 1 if thermal else 0, _dp(out))
    return out[: 4 * q].reshape(q, 4)


class Oracle:
    """CPU restatement of the solver (same interface shape as capi.Solver)."""

    def __init__(self, resolution, relaxation_time, periodic=(True, True, True),
                 velocity_set="D3Q19", storage="fp32", multiphase=False,
                 boundary_mode="halfway", precision="f64"):
        self._lib = load(precision)
        X, Y, Z = (int(r) for r in resolution)
        vid = VELOCITY_SETS.get(velocity_set, velocity_set)
        st = STORAGE.get(storage, storage)
        bm = {"halfway": 0, "fullway_compat": 1}.get(boundary_mode, boundary_mode)
        h = self._lib.orc_create(X, Y, Z, vid, st, 1 if multiphase else 0, bm,
                                 int(bool(periodic[0])), int(bool(periodic[1])), int(bool(periodic[2])),
                                 float(np.float32(relaxation_time)))
        if not h:
            raise ValueError("invalid oracle configuration")
        self._h = h
        self.resolution = (X, Y, Z)
        self.cells = X * Y * Z
        self.q = self._lib.orc_q(h)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_param(self, name, value):
        v = _f64(np.float32(np.atleast_1d(value)))
        assert self._lib.orc_set_param(self._h, PARAMS.get(name, name), _dp(v), v.size) == 0

    def set_object(self, oid, velocity=(0, 0, 0), temperature=1.0, effective_density=1.0, **_):
        v = _f64(np.float32(velocity), 3)
        assert self._lib.orc_set_object(self._h, int(oid), _dp(v), float(np.float32(temperature)),
                                        float(np.float32(effective_density))) == 0

    def init_fields(self, rho, u, force=None, temperature=None, boundary_id=None):
        n = self.cells
        f32 = lambda a, k: None if a is None else _f64(np.asarray(a, np.float32), k)  # noqa: E731
        rho, u, force, temperature = f32(rho, n), f32(u, 3 * n), f32(force, 3 * n), f32(temperature, n)
        ids = None
        if boundary_id is not None:
            ids = np.ascontiguousarray(boundary_id, dtype=np.uint8).ravel()
        self._keep = (rho, u, force, temperature, ids)
        self._lib.orc_init_fields(self._h, _dp(rho), _dp(u), _dp(force), _dp(temperature),
                                  None if ids is None else ids.ctypes.data_as(C.POINTER(C.c_uint8)))

    def step(self, n=1, save=False, conservation=False):
        self._lib.orc_step(self._h, int(n), (1 if save else 0) | (2 if conservation else 0))

    @property
    def ticks(self):
        return int(self._lib.orc_ticks(self._h))

    @property
    def thermal(self):
        return bool(self._lib.orc_thermal(self._h))

    @property
    def forcing(self):
        return int(self._lib.orc_forcing(self._h))

    def read(self, field):
        n = self.cells
        fid = {"velocity_density": 0, "force_temperature": 2, "populations_raw": 3,
               "thermal_raw": 4, "psi": 5}[field]
        size = {0: 4 * n, 2: 4 * n, 3: self.q * n, 4: self._lib.orc_qt(self._h) * n, 5: n}[fid]
        out = np.zeros(max(size, 1), np.float64)
        assert self._lib.orc_read_field(self._h, fid, _dp(out)) == 0, field
        if fid in (0, 2):
            return out.reshape(n, 4)
        if fid in (3, 4):
            return out.reshape(-1, n)
        return out

    def write_populations(self, f):
        f = _f64(f, self.q * self.cells)
        self._lib.orc_write_populations(self._h, _dp(f))

    def conservation(self):
        out = np.zeros(5, np.float64)
        self._lib.orc_conservation(self._h, _dp(out))
        return out

    def last_step_sums(self):
        out = np.zeros(5, np.float64)
        self._lib.orc_last_sums(self._h, _dp(out))
        return out

    def aa_maps(self, odd):
        k = self.q * self.cells
        rd = np.zeros(k, np.int64)
        wr = np.zeros(k, np.int64)
        self._lib.orc_aa_maps(self._h, int(odd), rd.ctypes.data_as(C.POINTER(C.c_long)),
                              wr.ctypes.data_as(C.POINTER(C.c_long)))
        return rd, wr

    def two_lattice(self, f, n):
        f = _f64(f, self.q * self.cells).copy()
        self._lib.orc_two_lattice(self._h, _dp(f), int(n))
        return f.reshape(self.q, self.cells)


class OracleChannel:
    """CPU restatement of the differentiable channel (bf_oracle_channel.c)."""

    def __init__(self, resolution, relaxation_time, u_in, velocity_set="D2Q9", rho_out=1.0, precision="f64"):
        self._lib = load(precision)
        X, Y, Z = (int(r) for r in resolution)
        vid = VELOCITY_SETS[velocity_set]
        h = self._lib.och_create(X, Y, Z, vid, float(np.float32(relaxation_time)), float(np.float32(u_in)),
                                 float(np.float32(rho_out)))
        if not h:
            raise ValueError("invalid channel configuration")
        self._h = h
        self.resolution = (X, Y, Z)
        self.cells = X * Y * Z
        self.q = 9 if vid == 2 else 19

    def close(self):
        if getattr(self, "_h", None):
            self._lib.och_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_porosity(self, gamma):
        g = _f64(np.asarray(gamma, np.float32), self.cells)
        self._lib.och_set_porosity(self._h, _dp(g))

    def set_inlet(self, u_in):
        self._lib.och_set_inlet(self._h, float(u_in))

    def profile(self):
        out = np.zeros(self.resolution[1])
        self._lib.och_profile(self._h, _dp(out))
        return out

    def forward(self, steps):
        return float(self._lib.och_forward(self._h, int(steps)))

    def moments(self):
        out = np.zeros(4 * self.cells)
        self._lib.och_moments(self._h, _dp(out))
        return out.reshape(self.cells, 4)

    def precollision(self, state, x, y, z=0):
        s = _f64(state, self.q * self.cells)
        out = np.zeros(self.q)
        self._lib.och_precollision(self._h, _dp(s), int(x), int(y), int(z), _dp(out))
        return out

    def gradient(self, steps):
        dg = np.zeros(self.cells)
        du = np.zeros(1)
        J = self._lib.och_gradient(self._h, int(steps), _dp(dg), _dp(du))
        return float(J), dg, float(du[0])

    def gradient_checkpointed(self, steps, checkpoint_every, half_states=False):
        """The device's schedule: checkpoints every K ticks, segments recomputed
        (their states kept as shifted FP16 when half_states)."""
        dg = np.zeros(self.cells)
        du = np.zeros(1)
        J = self._lib.och_gradient_checkpointed(self._h, int(steps), int(checkpoint_every), int(bool(half_states)),
                                                _dp(dg), _dp(du))
        return float(J), dg, float(du[0])

    def tick_pair(self, state, nu):
        s = _f64(state, self.q * self.cells)
        n = _f64(nu, self.q * self.cells)
        out = np.zeros_like(s)
        nup = np.zeros_like(s)
        self._lib.och_tick_pair(self._h, _dp(s), _dp(n), _dp(out), _dp(nup))
        return out, nup
