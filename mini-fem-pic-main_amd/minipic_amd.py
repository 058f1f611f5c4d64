"""ctypes binding of libminipic_amd.so (include/mp_abi.h, include/mp_mesh.h).

Thin plumbing for tests and the benchmark: numpy arrays in, numpy arrays out, every call
goes through the C ABI.  There is no fallback: if the shared library is missing or was
built for another target, loading fails loudly.

Reference call surface mirrored (AD1641/mini-fem-pic-main):
  DeviceData(volume)                     src/mini-pic.cpp:39-53, 322-338  -> Mover(mesh)
  MoveParticles(ions, volume, solver, params, device)
                                         src/mini-pic.cpp:268-320         -> Mover.move_particles
  LoadVolumeMesh / LoadSurfaceMesh       src/meshes.cpp:28-357            -> Mesh.load_dat / load_surface
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libminipic_amd.so")

MP_OK = 0
SCATTER_ATOMIC = 0
SCATTER_DETERMINISTIC = 1
TUNE_MULTI_STEP = 0x100  # OR into autotune's mode: time candidates through mp_step_many
WALL_ABSORB = 0
WALL_REFLECT = 1
NOALLOC_MAX_PARTICLES = 1 << 22  # MP_NOALLOC_MAX_PARTICLES
NODE_NORMAL, NODE_OPEN, NODE_INLET, NODE_FIXED = 0, 1, 2, 3

# physical constants of the reference (src/FESolver.h:18-21)
EPS0 = 8.8541878e-12
QE = 1.602e-19
AMU = 1.660538921e-27
KB = 8.617333262e-5

# Tetra, src/meshes.h:33-43 (168 B) and Particle, src/particles.h:20-25 (88 B)
TET_DTYPE = np.dtype([("con", "<i4", 4), ("volume", "<f8"), ("alpha", "<f8", 4), ("beta", "<f8", 4),
                      ("gamma", "<f8", 4), ("delta", "<f8", 4), ("cell_con", "<i4", 4)])
PARTICLE_DTYPE = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("lc", "<f8", 4),
                           ("cell_index", "<i4"), ("pad", "<i4")])
FACE_DTYPE = np.dtype([("con", "<i4", 3), ("area", "<f8"), ("u", "<f8", 3), ("v", "<f8", 3),
                       ("cell_con", "<i4"), ("normal", "<f8", 3)], align=True)
assert TET_DTYPE.itemsize == 168 and PARTICLE_DTYPE.itemsize == 88


class MpOptions(C.Structure):
    _fields_ = [("sort_interval", C.c_int32), ("wall_mode", C.c_int32), ("block_size", C.c_int32),
                ("particles_per_thread", C.c_int32), ("walk_cap", C.c_int32), ("scatter_agg", C.c_int32),
                ("sort_shift", C.c_int32), ("keep_pid_order", C.c_int32),
                ("gather_sort", C.c_int32), ("sort_order", C.c_int32), ("det_algo", C.c_int32),
                ("sort_pid_bits", C.c_int32)]


class MpTuneTrial(C.Structure):
    _fields_ = [("opt", MpOptions), ("ms_per_step", C.c_double)]


class MpFeStats(C.Structure):
    _fields_ = [("nr_iterations", C.c_int32), ("gs_sweeps", C.c_int32), ("gs_unconverged", C.c_int32),
                ("zero_diagonal", C.c_int32), ("nr_l2", C.c_double)]


class MpCounters(C.Structure):
    _fields_ = [("particle_steps", C.c_int64), ("lost", C.c_int64), ("crossed", C.c_int64),
                ("moves", C.c_int64), ("cap_hits", C.c_int64), ("reflections", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


# every exported symbol and its (restype, argtypes)
_P = C.c_void_p
_I32, _I64, _D = C.c_int32, C.c_int64, C.c_double
SIGNATURES = {
    # mp_abi.h
    "mp_create": (C.c_int, [_P, _I32, _P, _I32, C.POINTER(_P)]),
    "mp_destroy": (None, [_P]),
    "mp_last_error": (C.c_char_p, [_P]),
    "mp_get_options": (C.c_int, [_P, C.POINTER(MpOptions)]),
    "mp_set_options": (C.c_int, [_P, C.POINTER(MpOptions)]),
    "mp_set_sort_interval": (C.c_int, [_P, _I32]),
    "mp_upload_field": (C.c_int, [_P, _P, _I32]),
    "mp_step_aos": (C.c_int, [_P, _P, C.POINTER(C.c_size_t), _D, _D, _D, _P, _I32]),
    "mp_step_aos_noalloc": (C.c_int, [_P, _P, _I64, _P, C.POINTER(_I64), _D, _D, _D, _P, _I32]),
    "mp_set_particles": (C.c_int, [_P, _P, _P, _P, _I64]),
    "mp_add_particles": (C.c_int, [_P, _P, _P, _P, _I64]),
    "mp_step": (C.c_int, [_P, _D, _D, _D, _I32]),
    "mp_step_many": (C.c_int, [_P, _D, _D, _D, _I32, _I32]),
    "mp_get_density_step": (C.c_int, [_P, _I32, _P]),
    "mp_sync": (C.c_int, [_P]),
    "mp_num_active": (C.c_int, [_P, C.POINTER(_I64)]),
    "mp_get_slots": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.POINTER(_I64)]),
    "mp_get_tet_order": (C.c_int, [_P, C.c_void_p]),
    "mp_get_particles": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(_I64)]),
    "mp_get_density": (C.c_int, [_P, _P]),
    "mp_get_counters": (C.c_int, [_P, C.POINTER(MpCounters)]),
    "mp_reset_counters": (C.c_int, [_P]),
    "mp_get_det_paths": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mp_sort_now": (C.c_int, [_P]),
    "mp_stream": (_P, [_P]),
    "mp_set_timing": (C.c_int, [_P, _I32]),
    "mp_get_timing": (C.c_int, [_P, C.POINTER(_D), C.POINTER(_I64)]),
    "mp_adjoint": (C.c_int, [_P, _P, _P, _P, _I64, _D, _D, _D, _I32, _P, _P, _I32, C.POINTER(_D), _P,
                             C.POINTER(_D)]),
    "mp_adjoint_ex": (C.c_int, [_P, _P, _P, _P, _I64, _D, _D, _D, _I32, _P, _P, _I32, C.POINTER(_D), _P,
                                C.POINTER(_D), _I32]),
    "mp_autotune": (C.c_int, [_P, _D, _D, _D, _I32, _I32, C.POINTER(MpOptions), C.POINTER(_D),
                              C.POINTER(MpTuneTrial), _I32, C.POINTER(_I32)]),
    "mp_batch_create": (C.c_int, [_P, _I32, C.POINTER(_P)]),
    "mp_batch_destroy": (None, [_P]),
    "mp_batch_set_run": (C.c_int, [_P, _I32, _P, _P, _P, _I64, _D, _D, _D]),
    "mp_batch_step": (C.c_int, [_P, _I32]),
    "mp_batch_step_sequential": (C.c_int, [_P, _I32]),
    "mp_batch_step_many": (C.c_int, [_P, _I32, _I32]),
    "mp_batch_get_density_step": (C.c_int, [_P, _I32, _I32, _P]),
    "mp_batch_set_graph": (C.c_int, [_P, _I32]),
    "mp_batch_sync": (C.c_int, [_P]),
    "mp_batch_get_density": (C.c_int, [_P, _I32, _P]),
    "mp_batch_get_particles": (C.c_int, [_P, _I32, _P, _P, _P, _P, C.POINTER(_I64)]),
    "mp_batch_get_counters": (C.c_int, [_P, C.POINTER(MpCounters)]),
    "mp_batch_reset_counters": (C.c_int, [_P]),
    # mp_mesh.h
    "mp_mesh_box": (C.c_int, [_I32, _I32, _I32, _D, _D, _D, C.POINTER(_P)]),
    "mp_mesh_load_dat": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "mp_mesh_load_surface": (C.c_int, [_P, C.c_char_p, _I32, _I32]),
    "mp_mesh_free": (None, [_P]),
    "mp_mesh_num_nodes": (_I32, [_P]),
    "mp_mesh_num_tets": (_I32, [_P]),
    "mp_mesh_tets": (_P, [_P]),
    "mp_mesh_node_pos": (_P, [_P]),
    "mp_mesh_node_volume": (_P, [_P]),
    "mp_mesh_node_type": (_P, [_P]),
    "mp_mesh_last_error": (C.c_char_p, []),
    # mp_fem.h
    "mp_fe_create": (C.c_int, [_P, _D, _D, _D, _D, _I32, C.POINTER(_P)]),
    "mp_fe_free": (None, [_P]),
    "mp_fe_num_unknowns": (_I32, [_P]),
    "mp_fe_compute_phi": (C.c_int, [_P, _P, _I32, C.POINTER(MpFeStats)]),
    "mp_fe_update_ef": (C.c_int, [_P, _P]),
    "mp_fe_get_phi": (C.c_int, [_P, _P]),
    "mp_fe_get_stiffness": (C.c_int, [_P, C.POINTER(_I64), _P, _P, _P]),
    "mp_fe_get_profile": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(_I64),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "mp_fe_last_error": (C.c_char_p, []),
    # mp_output.h
    "mp_write_mesh_vtu": (C.c_int, [_P, C.c_char_p, _P, _P, _P]),
    "mp_write_particles_vtp": (C.c_int, [C.c_char_p, _P, _I64, _I64]),
    "mp_output_last_error": (C.c_char_p, []),
    # mp_inject.h
    "mp_rng_create": (C.c_int, [C.c_uint32, C.POINTER(_P)]),
    "mp_rng_free": (None, [_P]),
    "mp_rng_uniform": (C.c_int, [_P, _P, _I64]),
    "mp_inject_count": (C.c_int, [_P, _D, _D, _D, _D, _D, C.POINTER(_I64)]),
    "mp_inject_ions": (C.c_int, [_P, _P, _D, _D, _D, _D, _D, _P, C.POINTER(_D), _P, _I64, C.POINTER(_I64)]),
    "mp_inject_ions_thermal": (C.c_int, [_P, _P, _D, _D, _D, _D, _D, _D, _P, C.POINTER(_D), _P, _I64,
                                         C.POINTER(_I64)]),
    "mp_mesh_num_inlet_faces": (_I32, [_P]),
    "mp_mesh_inlet_faces": (_P, [_P]),
    "mp_mesh_locate_box": (C.c_int, [_P, _P, _I64, _P]),
    "mp_field_from_potential": (C.c_int, [_P, _P, _P]),
    "mp_field_sinusoid": (C.c_int, [_P, _D, _P]),
    "mp_synth_uniform": (C.c_int, [C.c_uint64, _I64, _P, _P, _D, _P, _P, _P]),
    "mp_synth_mixture": (C.c_int, [C.c_uint64, _I64, _P, _P, _I32, _P, _P, _D, _D, _P, _P, _P]),
}

_lib = None


def lib():
    """Load libminipic_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[mp error {code}] {msg}")
        self.code = code


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


# ---------------------------------------------------------------------------------------
class Mesh:
    """Tet mesh with the tables LoadVolumeMesh precomputes (host-side, no GPU needed)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        L = lib()
        self.n_nodes = int(L.mp_mesh_num_nodes(self._h))
        self.n_tets = int(L.mp_mesh_num_tets(self._h))
        self.box = None

    @classmethod
    def box_mesh(cls, n, side=1.0, ny=None, nz=None, ly=None, lz=None):
        ny = n if ny is None else ny
        nz = n if nz is None else nz
        ly = side if ly is None else ly
        lz = side if lz is None else lz
        h = C.c_void_p()
        rc = lib().mp_mesh_box(n, ny, nz, side, ly, lz, C.byref(h))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_mesh_last_error().decode())
        m = cls(h.value)
        m.box = (n, ny, nz, side, ly, lz)
        return m

    @classmethod
    def load_dat(cls, path):
        h = C.c_void_p()
        rc = lib().mp_mesh_load_dat(str(path).encode(), C.byref(h))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_mesh_last_error().decode())
        return cls(h.value)

    def load_surface(self, path, node_type, invert_normals=False):
        rc = lib().mp_mesh_load_surface(self._h, str(path).encode(), node_type, int(bool(invert_normals)))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_mesh_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.mp_mesh_free(self._h)
            self._h = C.c_void_p()

    def _view(self, fn, dtype, count, width=None):
        p = fn(self._h)
        n = count * (1 if width is None else width)
        buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(p)
        a = np.frombuffer(buf, dtype=dtype, count=n).copy()
        return a if width is None else a.reshape(count, width)

    @property
    def tets(self):
        return self._view(lib().mp_mesh_tets, TET_DTYPE, self.n_tets)

    @property
    def node_pos(self):
        return self._view(lib().mp_mesh_node_pos, np.float64, self.n_nodes, 3)

    @property
    def node_volume(self):
        return self._view(lib().mp_mesh_node_volume, np.float64, self.n_nodes)

    @property
    def node_type(self):
        return self._view(lib().mp_mesh_node_type, np.int32, self.n_nodes)

    @property
    def inlet_faces(self):
        n = int(lib().mp_mesh_num_inlet_faces(self._h))
        if n == 0:
            return np.zeros(0, FACE_DTYPE)
        return self._view(lib().mp_mesh_inlet_faces, FACE_DTYPE, n)

    def locate_box(self, pos):
        pos = _f64(pos, (-1, 3))
        cell = np.empty(len(pos), np.int32)
        rc = lib().mp_mesh_locate_box(self._h, _ptr(pos), len(pos), _ptr(cell))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_mesh_last_error().decode())
        return cell

    def field_from_potential(self, phi):
        phi = _f64(phi)
        assert phi.shape == (self.n_nodes,)
        ef = np.empty((self.n_tets, 3), np.float64)
        rc = lib().mp_field_from_potential(self._h, _ptr(phi), _ptr(ef))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_mesh_last_error().decode())
        return ef

    def field_sinusoid(self, amp=10.0):
        ef = np.empty((self.n_tets, 3), np.float64)
        rc = lib().mp_field_sinusoid(self._h, amp, _ptr(ef))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_mesh_last_error().decode())
        return ef


def synth_uniform(seed, n, lo, hi, sigma, drift):
    pos = np.empty((n, 3), np.float64)
    vel = np.empty((n, 3), np.float64)
    lo, hi, drift = _f64(lo), _f64(hi), _f64(drift)
    rc = lib().mp_synth_uniform(seed, n, _ptr(lo), _ptr(hi), sigma, _ptr(drift), _ptr(pos), _ptr(vel))
    if rc != MP_OK:
        raise MpError(rc, lib().mp_mesh_last_error().decode())
    return pos, vel


def synth_mixture(seed, n, lo, hi, weights, centres, sigma_pos, sigma_vel, drift):
    pos = np.empty((n, 3), np.float64)
    vel = np.empty((n, 3), np.float64)
    lo, hi, drift = _f64(lo), _f64(hi), _f64(drift)
    w, c = _f64(weights), _f64(centres, (-1, 3))
    rc = lib().mp_synth_mixture(seed, n, _ptr(lo), _ptr(hi), len(w), _ptr(w), _ptr(c), sigma_pos, sigma_vel,
                                _ptr(drift), _ptr(pos), _ptr(vel))
    if rc != MP_OK:
        raise MpError(rc, lib().mp_mesh_last_error().decode())
    return pos, vel


# ---------------------------------------------------------------------------------------
FE_NONLINEAR, FE_GAUSS_SEIDEL, FE_DIRECT = 0, 1, 2
FE_RESET_F1 = 1
FE_PROFILE_ENTRIES = 11
KB = 8.617333262e-5   # src/FESolver.h:21, eV/K
EPS0 = 8.8541878e-12  # src/FESolver.h:18


class FESolver:
    """Host FE Poisson-Boltzmann solver (FESolver, src/FESolver.cpp) over a Mesh's node types."""

    def __init__(self, mesh, n0, phi0, kTe, wall_potential=0.0, flags=0):
        h = C.c_void_p()
        rc = lib().mp_fe_create(mesh._h, n0, phi0, kTe, wall_potential, flags, C.byref(h))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_fe_last_error().decode())
        self._h = h
        self.mesh = mesh
        self.neq = int(lib().mp_fe_num_unknowns(h))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.mp_fe_free(self._h)
            self._h = C.c_void_p()

    def compute_phi(self, ion_den, method=FE_NONLINEAR):
        den = _f64(ion_den)
        assert den.shape == (self.mesh.n_nodes,)
        st = MpFeStats()
        rc = lib().mp_fe_compute_phi(self._h, _ptr(den), method, C.byref(st))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_fe_last_error().decode())
        return {k: (float if k == "nr_l2" else int)(getattr(st, k)) for k, _ in MpFeStats._fields_}

    def update_ef(self):
        ef = np.empty((self.mesh.n_tets, 3), np.float64)
        rc = lib().mp_fe_update_ef(self._h, _ptr(ef))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_fe_last_error().decode())
        return ef

    @property
    def phi(self):
        uh = np.empty(self.mesh.n_nodes, np.float64)
        rc = lib().mp_fe_get_phi(self._h, _ptr(uh))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_fe_last_error().decode())
        return uh

    def stiffness(self):
        """K over the unknowns as (row_ptr, col, val)."""
        nnz = C.c_int64()
        lib().mp_fe_get_stiffness(self._h, C.byref(nnz), None, None, None)
        rp = np.empty(self.neq + 1, np.int64)
        ci = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.float64)
        rc = lib().mp_fe_get_stiffness(self._h, C.byref(nnz), _ptr(rp), _ptr(ci), _ptr(val))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_fe_last_error().decode())
        return rp, ci, val

    def profile(self):
        """The solver's method timings since creation, as trace rows: {name: (calls, inclusive
        seconds, seconds in traced callees)} (FESolver's TRACE_ME rows, src/FESolver.cpp)."""
        rows = {}
        for k in range(FE_PROFILE_ENTRIES):
            name, calls = C.c_char_p(), C.c_int64()
            total, child = C.c_double(), C.c_double()
            rc = lib().mp_fe_get_profile(self._h, k, C.byref(name), C.byref(calls), C.byref(total), C.byref(child))
            if rc != MP_OK:
                raise MpError(rc, lib().mp_fe_last_error().decode())
            rows[name.value.decode()] = (int(calls.value), float(total.value), float(child.value))
        return rows


class Rng:
    """rnd() of src/maths.cpp:17-19: std::mt19937(seed) + uniform_real_distribution<double>(0, 1)."""

    def __init__(self, seed=0):
        h = C.c_void_p()
        rc = lib().mp_rng_create(seed, C.byref(h))
        if rc != MP_OK:
            raise MpError(rc, "mp_rng_create failed")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.mp_rng_free(self._h)
            self._h = C.c_void_p()

    def uniform(self, n):
        out = np.empty(n, np.float64)
        rc = lib().mp_rng_uniform(self._h, _ptr(out), n)
        if rc != MP_OK:
            raise MpError(rc, "mp_rng_uniform failed")
        return out


def write_mesh_vtu(mesh, path, phi, ef, ion_den):
    """OutputMesh (src/meshes.cpp:360-449) to path."""
    phi, ef, den = _f64(phi), _f64(ef, (-1, 3)), _f64(ion_den)
    rc = lib().mp_write_mesh_vtu(mesh._h, str(path).encode(), _ptr(phi), _ptr(ef), _ptr(den))
    if rc != MP_OK:
        raise MpError(rc, lib().mp_output_last_error().decode())


def write_particles_vtp(path, pos):
    """OutputParticles (src/particles.cpp:20-45): pos is [n][3] or Particle records."""
    if pos.dtype == PARTICLE_DTYPE:
        arr, stride = np.ascontiguousarray(pos), PARTICLE_DTYPE.itemsize
    else:
        arr, stride = _f64(pos, (-1, 3)), 24
    rc = lib().mp_write_particles_vtp(str(path).encode(), _ptr(arr), len(arr), stride)
    if rc != MP_OK:
        raise MpError(rc, lib().mp_output_last_error().decode())


def thermal_speed(temperature_k, mass):
    """Per-component standard deviation sqrt(kT/m) (m/s) of a Maxwellian at temperature_k
    kelvin for particles of mass kg (k = KB eV/K times QE J/eV, the reference's constants)."""
    return float(np.sqrt(KB * temperature_k * QE / mass))


def inject_ions(mesh, rng, plasma_den, ion_velocity, dt, spwt, q_over_m, ef, rem, v_thermal=0.0):
    """InjectIons (src/mini-pic.cpp:208-265): (new Particle records, carried fraction).
    v_thermal > 0 adds a Maxwellian spread of that per-component deviation (m/s) to the cold
    beam (mp_inject_ions_thermal; the reference's own injection is v_thermal = 0)."""
    n = C.c_int64()
    rc = lib().mp_inject_count(mesh._h, plasma_den, ion_velocity, dt, spwt, rem, C.byref(n))
    if rc != MP_OK:
        raise MpError(rc, "mp_inject_count failed")
    out = np.zeros(max(n.value, 1), PARTICLE_DTYPE)
    ef = _f64(ef, (-1, 3))
    r = C.c_double(rem)
    got = C.c_int64()
    rc = lib().mp_inject_ions_thermal(mesh._h, rng._h, plasma_den, ion_velocity, v_thermal, dt, spwt, q_over_m,
                                      _ptr(ef), C.byref(r), _ptr(out), len(out), C.byref(got))
    if rc != MP_OK:
        raise MpError(rc, "mp_inject_ions failed")
    return out[:got.value], float(r.value)


# ---------------------------------------------------------------------------------------
class Mover:
    """One device context (DeviceData): the Tetra table on the GPU plus particle buffers."""

    def __init__(self, tets, node_volume):
        tets = np.ascontiguousarray(tets, dtype=TET_DTYPE)
        node_volume = _f64(node_volume)
        self.n_tets = len(tets)
        self.n_nodes = len(node_volume)
        h = C.c_void_p()
        rc = lib().mp_create(_ptr(tets), self.n_tets, _ptr(node_volume), self.n_nodes, C.byref(h))
        if rc != MP_OK:
            raise MpError(rc, lib().mp_last_error(None).decode())
        self._h = h

    @classmethod
    def from_mesh(cls, mesh):
        return cls(mesh.tets, mesh.node_volume)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().mp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != MP_OK:
            raise MpError(rc, lib().mp_last_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    @property
    def options(self):
        o = MpOptions()
        self._check(lib().mp_get_options(self._h, C.byref(o)))
        return {k: int(getattr(o, k)) for k, _ in MpOptions._fields_}

    def set_options(self, **kw):
        o = MpOptions()
        self._check(lib().mp_get_options(self._h, C.byref(o)))
        for k, v in kw.items():
            setattr(o, k, int(v))
        self._check(lib().mp_set_options(self._h, C.byref(o)))

    def upload_field(self, ef):
        ef = _f64(ef, (-1, 3))
        self._check(lib().mp_upload_field(self._h, _ptr(ef), len(ef)))

    # drop-in MoveParticles over the reference Particle array (in place; returns kept count)
    def step_aos(self, parts, q_over_m, dt, spwt, mode=SCATTER_ATOMIC):
        assert parts.dtype == PARTICLE_DTYPE and parts.flags.c_contiguous
        n = C.c_size_t(len(parts))
        den = np.empty(self.n_nodes, np.float64)
        self._check(lib().mp_step_aos(self._h, _ptr(parts), C.byref(n), q_over_m, dt, spwt, _ptr(den), mode))
        return int(n.value), den

    def step_aos_noalloc(self, parts_dev, n, kept_dev, q_over_m, dt, spwt, den_dev, mode=SCATTER_ATOMIC):
        """mp_step_aos_noalloc over device addresses (ints): parts_dev holds n Particle records,
        kept_dev receives the survivors, den_dev the n_nodes normalised densities.  Returns the
        number kept."""
        k = C.c_int64()
        self._check(lib().mp_step_aos_noalloc(self._h, C.c_void_p(parts_dev), n, C.c_void_p(kept_dev), C.byref(k),
                                              q_over_m, dt, spwt, C.c_void_p(den_dev), mode))
        return int(k.value)

    def move_particles(self, parts, q_over_m, dt, spwt, mode=SCATTER_ATOMIC):
        """MoveParticles semantics on a numpy Particle array: returns (kept particles, den)."""
        parts = np.ascontiguousarray(parts, dtype=PARTICLE_DTYPE).copy()
        k, den = self.step_aos(parts, q_over_m, dt, spwt, mode)
        return parts[:k].copy(), den

    # device-resident path
    def set_particles(self, pos, vel, cell):
        pos, vel = _f64(pos, (-1, 3)), _f64(vel, (-1, 3))
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        assert len(pos) == len(vel) == len(cell)
        self._check(lib().mp_set_particles(self._h, _ptr(pos), _ptr(vel), _ptr(cell), len(pos)))

    def add_particles(self, pos, vel, cell):
        pos, vel = _f64(pos, (-1, 3)), _f64(vel, (-1, 3))
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        self._check(lib().mp_add_particles(self._h, _ptr(pos), _ptr(vel), _ptr(cell), len(pos)))

    def step(self, q_over_m, dt, spwt, mode=SCATTER_ATOMIC):
        self._check(lib().mp_step(self._h, q_over_m, dt, spwt, mode))

    def step_many(self, q_over_m, dt, spwt, mode=SCATTER_ATOMIC, n_steps=1):
        """n_steps steps under the frozen field; atomic mode keeps the particles in registers
        between the steps of a sort interval (one launch per interval)."""
        self._check(lib().mp_step_many(self._h, q_over_m, dt, spwt, mode, n_steps))

    def density_step(self, k):
        den = np.empty(self.n_nodes, np.float64)
        self._check(lib().mp_get_density_step(self._h, k, _ptr(den)))
        return den

    def sync(self):
        self._check(lib().mp_sync(self._h))

    def sort_now(self):
        self._check(lib().mp_sort_now(self._h))

    def num_active(self):
        n = C.c_int64()
        self._check(lib().mp_num_active(self._h, C.byref(n)))
        return int(n.value)

    def slots(self):
        """The resident slots in their current order (bucket order after a cell sort), lost
        ones included until the next sort: (cell, pid), int32 each."""
        n = C.c_int64(0)
        self._check(lib().mp_get_slots(self._h, None, None, C.byref(n)))
        cell, pid = np.empty(n.value, np.int32), np.empty(n.value, np.int32)
        self._check(lib().mp_get_slots(self._h, _ptr(cell), _ptr(pid), C.byref(n)))
        return cell, pid

    def tet_order(self, n_tets):
        """Rank of every tet along the Morton curve of sort_order 1 (int32, a permutation)."""
        rank = np.empty(n_tets, np.int32)
        self._check(lib().mp_get_tet_order(self._h, _ptr(rank)))
        return rank

    def get_particles(self):
        """Particles still in the domain, ascending pid: (pos, vel, cell, pid)."""
        n = C.c_int64(0)
        rc = lib().mp_get_particles(self._h, None, None, None, None, C.byref(n))
        cap = n.value
        if rc != MP_OK and cap == 0:
            self._check(rc)
        pos = np.empty((cap, 3), np.float64)
        vel = np.empty((cap, 3), np.float64)
        cell = np.empty(cap, np.int32)
        pid = np.empty(cap, np.int64)
        n = C.c_int64(cap)
        self._check(lib().mp_get_particles(self._h, _ptr(pos), _ptr(vel), _ptr(cell), _ptr(pid), C.byref(n)))
        k = n.value
        return pos[:k], vel[:k], cell[:k], pid[:k]

    def density(self):
        den = np.empty(self.n_nodes, np.float64)
        self._check(lib().mp_get_density(self._h, _ptr(den)))
        return den

    def counters(self):
        c = MpCounters()
        self._check(lib().mp_get_counters(self._h, C.byref(c)))
        return c.as_dict()

    def reset_counters(self):
        self._check(lib().mp_reset_counters(self._h))

    def det_paths(self):
        """(placed, rebuilt): deterministic steps whose records the push placed into slab segments
        (det_algo 2), and steps whose records were rebuilt from the pushed state."""
        a, b = C.c_int64(), C.c_int64()
        self._check(lib().mp_get_det_paths(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stream(self):
        return lib().mp_stream(self._h)

    def adjoint(self, pos0, vel0, cell0, q_over_m, dt, spwt, steps, den_target, drift_dir, checkpoint_every=10,
                mode=SCATTER_ATOMIC):
        """(loss, dL/dE [n_tets][3], dL/d drift) of K frozen-field steps (mp_adjoint_ex; mode
        SCATTER_DETERMINISTIC gives the same bits on every run)."""
        pos0, vel0 = _f64(pos0, (-1, 3)), _f64(vel0, (-1, 3))
        cell0 = np.ascontiguousarray(cell0, dtype=np.int32)
        tgt, dd = _f64(den_target), _f64(drift_dir)
        loss, dd_out = C.c_double(), C.c_double()
        g = np.empty((self.n_tets, 3), np.float64)
        self._check(lib().mp_adjoint_ex(self._h, _ptr(pos0), _ptr(vel0), _ptr(cell0), len(pos0), q_over_m, dt,
                                        spwt, steps, _ptr(tgt), _ptr(dd), checkpoint_every, C.byref(loss), _ptr(g),
                                        C.byref(dd_out), mode))
        return float(loss.value), g, float(dd_out.value)

    def autotune(self, q_over_m, dt, spwt, mode=SCATTER_ATOMIC, min_steps=10, max_trials=64, multi_step=False):
        """mp_autotune: (chosen options, their ms per step, [(options, ms per step)] of every
        timed candidate).  The chosen options are set; the particle state is left unchanged.
        multi_step: time the candidates through mp_step_many."""
        best, ms = MpOptions(), C.c_double()
        trials = (MpTuneTrial * max_trials)()
        nt = C.c_int32()
        m = mode | (TUNE_MULTI_STEP if multi_step else 0)
        self._check(lib().mp_autotune(self._h, q_over_m, dt, spwt, m, min_steps, C.byref(best), C.byref(ms),
                                      trials, max_trials, C.byref(nt)))
        as_dict = lambda o: {k: int(getattr(o, k)) for k, _ in MpOptions._fields_}  # noqa: E731
        log = [(as_dict(trials[k].opt), float(trials[k].ms_per_step)) for k in range(min(nt.value, max_trials))]
        return as_dict(best), float(ms.value), log

    def set_timing(self, enable=True):
        self._check(lib().mp_set_timing(self._h, int(bool(enable))))

    def timing(self):
        """(summed push-kernel ms, number of push launches) since set_timing(True)."""
        ms, n = C.c_double(), C.c_int64()
        self._check(lib().mp_get_timing(self._h, C.byref(ms), C.byref(n)))
        return float(ms.value), int(n.value)


class Batch:
    """Ensemble of independent runs on one Mover's mesh and field (mp_batch_*)."""

    def __init__(self, mover, n_runs):
        self.mover = mover
        self.n_runs = n_runs
        self.n_nodes = mover.n_nodes
        h = C.c_void_p()
        mover._check(lib().mp_batch_create(mover.handle, n_runs, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().mp_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_run(self, run, pos, vel, cell, q_over_m, dt, spwt):
        pos, vel = _f64(pos, (-1, 3)), _f64(vel, (-1, 3))
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        self.mover._check(lib().mp_batch_set_run(self._h, run, _ptr(pos), _ptr(vel), _ptr(cell), len(pos),
                                                 q_over_m, dt, spwt))

    def step(self, mode=SCATTER_ATOMIC):
        self.mover._check(lib().mp_batch_step(self._h, mode))

    def step_sequential(self, mode=SCATTER_ATOMIC):
        self.mover._check(lib().mp_batch_step_sequential(self._h, mode))

    def step_many(self, n_steps, mode=SCATTER_ATOMIC):
        """n_steps steps of every run; atomic mode keeps the particles in registers between the
        steps of a sort interval (one launch per interval)."""
        self.mover._check(lib().mp_batch_step_many(self._h, mode, n_steps))

    def density_step(self, k, run):
        den = np.empty(self.n_nodes, np.float64)
        self.mover._check(lib().mp_batch_get_density_step(self._h, k, run, _ptr(den)))
        return den

    def set_graph(self, enable=True):
        self.mover._check(lib().mp_batch_set_graph(self._h, int(bool(enable))))

    def sync(self):
        self.mover._check(lib().mp_batch_sync(self._h))

    def density(self, run):
        den = np.empty(self.n_nodes, np.float64)
        self.mover._check(lib().mp_batch_get_density(self._h, run, _ptr(den)))
        return den

    def get_particles(self, run):
        n = C.c_int64(0)
        lib().mp_batch_get_particles(self._h, run, None, None, None, None, C.byref(n))
        cap = n.value
        pos, vel = np.empty((cap, 3)), np.empty((cap, 3))
        cell, pid = np.empty(cap, np.int32), np.empty(cap, np.int64)
        n = C.c_int64(cap)
        self.mover._check(lib().mp_batch_get_particles(self._h, run, _ptr(pos), _ptr(vel), _ptr(cell), _ptr(pid),
                                                       C.byref(n)))
        k = n.value
        return pos[:k], vel[:k], cell[:k], pid[:k]

    def counters(self):
        c = MpCounters()
        self.mover._check(lib().mp_batch_get_counters(self._h, C.byref(c)))
        return c.as_dict()

    def reset_counters(self):
        self.mover._check(lib().mp_batch_reset_counters(self._h))


def particles_from_arrays(pos, vel, cell):
    """Reference Particle records (lc zeroed) from [n][3] arrays and cells."""
    p = np.zeros(len(pos), PARTICLE_DTYPE)
    p["pos"] = pos
    p["vel"] = vel
    p["cell_index"] = cell
    return p
