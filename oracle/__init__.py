"""CPU oracle of MoveParticles (TEST INFRASTRUCTURE ONLY).

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this package:
it is the checker, never the thing measured or shipped.

  liboracle.so      plain-C serial restatement of the mover (mp_oracle.c) and of the
                    reference's dense FE solver (fe_dense.c), -ffp-contract=off
  _ref/libref.so    the reference's own src/meshes.cpp, maths.cpp, trace.cpp, FESolver.cpp and
                    particles.cpp compiled where they lie, plus ref_harness.cpp (built only
                    where the reference exists)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libref.so")

TET_DTYPE = np.dtype([("con", "<i4", 4), ("volume", "<f8"), ("alpha", "<f8", 4), ("beta", "<f8", 4),
                      ("gamma", "<f8", 4), ("delta", "<f8", 4), ("cell_con", "<i4", 4)])
PARTICLE_DTYPE = np.dtype([("pos", "<f8", 3), ("vel", "<f8", 3), ("lc", "<f8", 4),
                           ("cell_index", "<i4"), ("pad", "<i4")])

_P = C.c_void_p
_I32, _I64, _D = C.c_int32, C.c_int64, C.c_double
_SIG = {
    "or_move_aos": (_I64, [_P, _P, _P, _I64, _D, _D, _P, _I32, _P]),
    "or_move": (_I64, [_P, _P, _P, _P, _P, _P, _I64, _D, _D, _P, _I32, _I32, _P]),
    "or_move_omp": (_I64, [_P, _P, _P, _P, _P, _I64, _D, _D, _P, _I32, _I32, _P]),
    "or_compact": (_I64, [_P, _P, _P, _P, _P, _I64]),
    "or_compact_aos": (_I64, [_P, _I64]),
    "or_normalise": (None, [_P, _P, _I32, _D]),
    "or_adjoint": (C.c_int, [_P, _P, _I32, _P, _I32, _P, _P, _P, _I64, _D, _D, _D, _I32, _P, _P, _I32, _I32,
                             _P, _P, _P]),
    "or_fe_create": (_P, [_P, _P, _I32, _P, _I32, _D, _D, _D, _D, C.c_int]),
    "or_fe_free": (None, [_P]),
    "or_fe_compute_phi": (C.c_int, [_P, _P, C.c_int]),
    "or_fe_get": (None, [_P, _P, _P]),
    "or_fe_neq": (_I32, [_P]),
}
_REF_SIG = {
    "ref_load_volume": (C.c_int, [C.c_char_p]),
    "ref_load_surface": (C.c_int, [C.c_char_p, C.c_int, C.c_int]),
    "ref_num_nodes": (_I32, []),
    "ref_num_tets": (_I32, []),
    "ref_num_inlet_faces": (_I32, []),
    "ref_get_tables": (None, [_P, _P, _P, _P]),
    "ref_get_inlet": (None, [_P, _P, _P, _P, _P, _P]),
    "ref_move_serial": (None, [_P, _P, _P, _I64, _D, _D, _P]),
    "ref_rnd": (None, [_I64, _P]),
    "ref_det3": (_D, [_P]),
    "ref_det4": (_D, [_P]),
    "ref_fe_create": (C.c_int, [_D, _D, _D, _D]),
    "ref_fe_compute_phi": (C.c_int, [_P, C.c_int]),
    "ref_fe_lapack_system": (C.c_int, [_P, _P, _P]),
    "ref_fe_get": (None, [_P, _P]),
    "ref_output_mesh": (None, [C.c_int, _P, _P, _P]),
    "ref_output_particles": (None, [_P, _I64]),
}

_lib = None
_ref = None


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run make -C oracle")
        L = C.CDLL(ORACLE_SO)
        for k, (r, a) in _SIG.items():
            getattr(L, k).restype = r
            getattr(L, k).argtypes = a
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        for k, (r, a) in _REF_SIG.items():
            getattr(L, k).restype = r
            getattr(L, k).argtypes = a
        _ref = L
    return _ref


def _stats():
    return np.zeros(6, np.int64)


STAT_KEYS = ("particle_steps", "lost", "crossed", "moves", "cap_hits", "reflections")


def move_aos(tets, ef, parts, qm, dt, den=None, walk_cap=4096):
    """Serial reference kernel over Particle records (in place). Returns (den_raw, stats)."""
    tets = np.ascontiguousarray(tets, dtype=TET_DTYPE)
    ef = np.ascontiguousarray(ef, dtype=np.float64)
    assert parts.dtype == PARTICLE_DTYPE and parts.flags.c_contiguous
    nn = int(tets["con"].max()) + 1 if den is None else len(den)
    if den is None:
        den = np.zeros(nn, np.float64)
    st = _stats()
    lib().or_move_aos(_ptr(tets), _ptr(ef), _ptr(parts), len(parts), qm, dt, _ptr(den), walk_cap, _ptr(st))
    return den, dict(zip(STAT_KEYS, st.tolist()))


def move(tets, ef, pos, vel, cell, qm, dt, n_nodes, wall_mode=0, walk_cap=4096, lc_out=False):
    """Serial pass over [n][3] arrays (in place); cell < 0 skipped. Returns (den_raw, stats, lc)."""
    tets = np.ascontiguousarray(tets, dtype=TET_DTYPE)
    ef = np.ascontiguousarray(ef, dtype=np.float64)
    for a in (pos, vel):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    assert cell.dtype == np.int32 and cell.flags.c_contiguous
    den = np.zeros(n_nodes, np.float64)
    lc = np.zeros((len(pos), 4), np.float64) if lc_out else None
    st = _stats()
    lib().or_move(_ptr(tets), _ptr(ef), _ptr(pos), _ptr(vel), _ptr(cell), _ptr(lc), len(pos), qm, dt,
                  _ptr(den), wall_mode, walk_cap, _ptr(st))
    return den, dict(zip(STAT_KEYS, st.tolist())), lc


def move_omp(tets, ef, pos, vel, cell, qm, dt, n_nodes, wall_mode=0, walk_cap=4096, den=None):
    tets = np.ascontiguousarray(tets, dtype=TET_DTYPE)
    if den is None:
        den = np.zeros(n_nodes, np.float64)
    thr = C.c_int32(0)
    n = lib().or_move_omp(_ptr(tets), _ptr(ef), _ptr(pos), _ptr(vel), _ptr(cell), len(pos), qm, dt, _ptr(den),
                          wall_mode, walk_cap, C.byref(thr))
    return den, int(n), int(thr.value)


def normalise(den, vol, spwt):
    den = np.ascontiguousarray(den, dtype=np.float64).copy()
    lib().or_normalise(_ptr(den), _ptr(np.ascontiguousarray(vol, dtype=np.float64)), len(den), spwt)
    return den


def compact(pos, vel, cell, pid=None):
    pid = np.arange(len(cell), dtype=np.int64) if pid is None else np.ascontiguousarray(pid, np.int64)
    k = lib().or_compact(_ptr(pos), _ptr(vel), _ptr(cell), None, _ptr(pid), len(cell))
    return pos[:k], vel[:k], cell[:k], pid[:k]


def compact_aos(parts):
    k = lib().or_compact_aos(_ptr(parts), len(parts))
    return parts[:k]


class DenseFE:
    """FESolver with the reference's dense K and J (fe_dense.c): the checker of the library's
    compressed-row solver.  node_type as in src/meshes.h:22; method 0 NonLinear, 1 GaussSeidel."""

    def __init__(self, node_pos, node_type, tets, n0, phi0, kTe, wall_potential=0.0, reset_f1=False):
        self.pos = np.ascontiguousarray(node_pos, np.float64)
        self.type = np.ascontiguousarray(node_type, np.int32)
        self.con = np.ascontiguousarray(np.asarray(tets)["con"], np.int32)
        self.nn, self.ne = len(self.type), len(self.con)
        self._h = lib().or_fe_create(_ptr(self.pos), _ptr(self.type), self.nn, _ptr(self.con), self.ne, n0, phi0,
                                     kTe, wall_potential, int(bool(reset_f1)))
        self.neq = int(lib().or_fe_neq(self._h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.or_fe_free(self._h)
            self._h = None

    def compute_phi(self, ion_den, method=0):
        den = np.ascontiguousarray(ion_den, np.float64)
        return int(lib().or_fe_compute_phi(self._h, _ptr(den), method))

    def phi_ef(self):
        uh = np.empty(self.nn)
        ef = np.empty((self.ne, 3))
        lib().or_fe_get(self._h, _ptr(uh), _ptr(ef))
        return uh, ef


class RefFE:
    """The reference's own FESolver (oracle/_ref) on the volume last loaded with
    ref().ref_load_volume / ref_load_surface, set up as the reference's main does.  Methods 0
    NonLinear and 1 GaussSeidel; the reference exits the process when a Newton solve does not
    converge."""

    def __init__(self, n_nodes, n_tets, n0, phi0, kTe, wall_potential=0.0):
        self.nn, self.ne = n_nodes, n_tets
        self.neq = int(ref().ref_fe_create(n0, phi0, kTe, wall_potential))
        if self.neq < 0:
            raise RuntimeError("ref_fe_create: no volume loaded")

    def compute_phi(self, ion_den, method=0):
        den = np.ascontiguousarray(ion_den, np.float64)
        if ref().ref_fe_compute_phi(_ptr(den), method) != 0:
            raise ValueError("method must be 0 (NonLinear) or 1 (GaussSeidel)")

    def lapack_system(self, ion_den):
        """(J, G) of the Lapack method's linearised step: J row-major with only j >= i set."""
        den = np.ascontiguousarray(ion_den, np.float64)
        J = np.zeros((self.neq, self.neq))
        G = np.zeros(self.neq)
        ref().ref_fe_lapack_system(_ptr(den), _ptr(J), _ptr(G))
        return J, G

    def phi_ef(self):
        uh = np.empty(self.nn)
        ef = np.empty((self.ne, 3))
        ref().ref_fe_get(_ptr(uh), _ptr(ef))
        return uh, ef


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def splitmix_draws(seed, items, j):
    """Draw j of each item under seed: mix64(seed + (16 item + j + 1) * 0x9E3779B97F4A7C15) >> 11,
    times 2^-53 (SplitMix64's output function on a Weyl sequence; uint64 arithmetic wraps)."""
    items = np.asarray(items, np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (items * np.uint64(16) + np.uint64(j + 1)) * np.uint64(0x9E3779B97F4A7C15)
        return (_mix64(z) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def thermal_normals(rng_seed, k0, n):
    """z[n][3] of the injected particles k0 .. k0+n-1 of a stream created with rng_seed: Box-Muller
    on draws (0, 1) and (2, 3) of item k under s = mix64(rng_seed ^ 0x7468726D616C0000), the
    third normal being the first of the second pair."""
    with np.errstate(over="ignore"):
        s = int(_mix64(np.uint64(rng_seed) ^ np.uint64(0x7468726D616C0000)))
    k = np.arange(k0, k0 + n, dtype=np.uint64)
    u = [splitmix_draws(s, k, j) for j in range(4)]
    r1, a1 = np.sqrt(-2.0 * np.log(1.0 - u[0])), 2.0 * np.pi * u[1]
    r2, a2 = np.sqrt(-2.0 * np.log(1.0 - u[2])), 2.0 * np.pi * u[3]
    return np.stack([r1 * np.cos(a1), r1 * np.sin(a1), r2 * np.cos(a2)], axis=1)


class AosMover:
    """MoveParticles (src/mini-pic.cpp:268-320) serially on the host: the reference kernel body
    over the Particle records, stable erase of the lost ones, den = raw * (spwt / node volume).
    Same call shape as the GPU mover of the coupled loop: (parts, ef) -> (kept parts, den)."""

    def __init__(self, tets, node_volume, qm, dt, spwt):
        self.tets = np.ascontiguousarray(tets, dtype=TET_DTYPE)
        self.vol = np.ascontiguousarray(node_volume, dtype=np.float64)
        self.qm, self.dt, self.spwt = qm, dt, spwt

    def __call__(self, parts, ef):
        parts = np.ascontiguousarray(parts)
        raw, _ = move_aos(self.tets, ef, parts, self.qm, self.dt, np.zeros(len(self.vol)))
        parts = compact_aos(parts).copy()
        return parts, normalise(raw, self.vol, self.spwt)


class Run:
    """Multi-step serial run over a particle set, tracking pid (stable erase each step)."""

    def __init__(self, tets, ef, node_volume, pos, vel, cell, wall_mode=0, walk_cap=4096):
        self.tets = np.ascontiguousarray(tets, dtype=TET_DTYPE)
        self.ef = np.ascontiguousarray(ef, dtype=np.float64)
        self.vol = np.ascontiguousarray(node_volume, dtype=np.float64)
        self.pos = np.array(pos, np.float64, order="C")
        self.vel = np.array(vel, np.float64, order="C")
        self.cell = np.array(cell, np.int32, order="C")
        self.pid = np.arange(len(self.cell), dtype=np.int64)
        self.wall_mode = wall_mode
        self.walk_cap = walk_cap
        self.stats = dict.fromkeys(STAT_KEYS, 0)
        self.den = np.zeros(len(self.vol))

    def step(self, qm, dt, spwt):
        raw, st, _ = move(self.tets, self.ef, self.pos, self.vel, self.cell, qm, dt, len(self.vol),
                          self.wall_mode, self.walk_cap)
        for k in STAT_KEYS:
            self.stats[k] += st[k]
        self.pos, self.vel, self.cell, self.pid = [np.ascontiguousarray(a) for a in
                                                   compact(self.pos, self.vel, self.cell, self.pid)]
        self.den = normalise(raw, self.vol, spwt)
        return self.den


def adjoint(tets, ef, vol, pos0, vel0, cell0, qm, dt, spwt, steps, den_target, drift_dir, walk_cap=4096,
            wall_mode=0):
    tets = np.ascontiguousarray(tets, dtype=TET_DTYPE)
    ef = np.ascontiguousarray(ef, dtype=np.float64)
    vol = np.ascontiguousarray(vol, dtype=np.float64)
    pos0 = np.ascontiguousarray(pos0, dtype=np.float64)
    vel0 = np.ascontiguousarray(vel0, dtype=np.float64)
    cell0 = np.ascontiguousarray(cell0, dtype=np.int32)
    tgt = np.ascontiguousarray(den_target, dtype=np.float64)
    dd = np.ascontiguousarray(drift_dir, dtype=np.float64)
    loss = C.c_double()
    dldd = C.c_double()
    g = np.zeros((len(tets), 3), np.float64)
    rc = lib().or_adjoint(_ptr(tets), _ptr(ef), len(tets), _ptr(vol), len(vol), _ptr(pos0), _ptr(vel0), _ptr(cell0),
                          len(cell0), qm, dt, spwt, steps, _ptr(tgt), _ptr(dd), wall_mode, walk_cap, C.byref(loss),
                          _ptr(g), C.byref(dldd))
    if rc != 0:
        raise RuntimeError(f"or_adjoint failed: {rc}")
    return float(loss.value), g, float(dldd.value)
