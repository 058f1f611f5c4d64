"""mp_step_aos_noalloc (the drop-in step over device-resident Particle records, every scratch
buffer in the library's fixed device array) against mp_step_aos (the same step with buffers the
library allocates) through the C ABI: two calls in a row, the second with fewer particles, so
the second call finds the first call's flags, prefix sums and tallies in the shared array.

Bar: kept particles (pos, vel, lc, cell) bit-exact and in the same order; density within
1e-12 relative (atomic scatter on both sides)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)
SPWT = 200.0
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    """The HIP runtime the mover library loaded (the caller's side allocates the buffers), not
    the copy a torch wheel bundles when an earlier test imported torch."""
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            p = line.split()[-1]
            if "libamdhip64" in p and "-packages/" not in p:
                path = p
                break
    assert path, "libamdhip64 not loaded"
    L = C.CDLL(path)
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0

    def put(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def test_noalloc_step_matches_step_aos(mp):
    mesh = mp.Mesh.box_mesh(8, 0.1)
    ef = mesh.field_sinusoid(10.0)
    dt = 4e-7
    sets = []
    for seed, n in ((3, 60000), (4, 25000)):
        pos, vel = mp.synth_uniform(seed, n, [0, 0, 0], [0.1] * 3, 2000.0, [5000.0, 0, 0])
        sets.append(mp.particles_from_arrays(pos, vel, mesh.locate_box(pos)))
    with mp.Mover.from_mesh(mesh) as mv:
        mv.upload_field(ef)
        hip = _hip()
        nmax = max(len(p) for p in sets)
        rec = mp.PARTICLE_DTYPE.itemsize
        parts_d, kept_d, den_d = DevBuf(hip, nmax * rec), DevBuf(hip, nmax * rec), DevBuf(hip, 8 * mesh.n_nodes)
        try:
            for call, parts in enumerate(sets):
                ref = parts.copy()
                k_ref, den_ref = mv.step_aos(ref, QM, dt, SPWT, mp.SCATTER_ATOMIC)
                parts_d.put(parts)
                k = mv.step_aos_noalloc(parts_d.p.value, len(parts), kept_d.p.value, QM, dt, SPWT, den_d.p.value)
                assert k == k_ref and 0 < k < len(parts), f"call {call}: kept count"
                got = kept_d.get(np.empty(k, mp.PARTICLE_DTYPE))
                assert got.tobytes() == ref[:k].tobytes(), f"call {call}: kept particles differ"
                den = den_d.get(np.empty(mesh.n_nodes))
                np.testing.assert_allclose(den, den_ref, rtol=1e-12, atol=0, err_msg=f"call {call}")
                assert den.sum() > 0
            # an empty call zeroes the density and keeps nothing
            assert mv.step_aos_noalloc(parts_d.p.value, 0, kept_d.p.value, QM, dt, SPWT, den_d.p.value) == 0
            assert not den_d.get(np.empty(mesh.n_nodes)).any()
            # sizes beyond the fixed scratch, the deterministic scatter, overlapping buffers and a bad
            # cell are refused before anything runs
            with pytest.raises(mp.MpError, match="at most"):
                mv.step_aos_noalloc(parts_d.p.value, mp.NOALLOC_MAX_PARTICLES + 1, kept_d.p.value, QM, dt, SPWT,
                                    den_d.p.value)
            with pytest.raises(mp.MpError, match="atomic scatter only"):
                mv.step_aos_noalloc(parts_d.p.value, 10, kept_d.p.value, QM, dt, SPWT, den_d.p.value,
                                    mp.SCATTER_DETERMINISTIC)
            with pytest.raises(mp.MpError, match="overlap"):
                mv.step_aos_noalloc(parts_d.p.value, 10, parts_d.p.value + rec, QM, dt, SPWT, den_d.p.value)
            bad = sets[1][:100].copy()
            bad["cell_index"][37] = mesh.n_tets
            parts_d.put(bad)
            with pytest.raises(mp.MpError, match="particle 37 has cell_index out of range"):
                mv.step_aos_noalloc(parts_d.p.value, 100, kept_d.p.value, QM, dt, SPWT, den_d.p.value)
        finally:
            for b in (parts_d, kept_d, den_d):
                b.free()
