"""Pins the CPU oracle and the host table code (no GPU).

The reference ships no tests or golden vectors, so the oracle is pinned by:
  * the reference's own code: src/meshes.cpp LoadVolumeMesh and src/meshes.h XtoLtet,
    compiled from the reference sources into oracle/_ref (skipped where absent), and the
    fixtures of tests/golden/ that make_golden.py wrote from it;
  * closed forms and brute force: exact push arithmetic, Kuhn-box volumes and barycentric
    identities, containing-tet search over all tets, single-particle scatter, closed-form
    reflection, and central finite differences for the adjoint.
"""
import gzip
import hashlib
import json
import os
import shutil
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN, REF_EXAMPLES

QE, AMU = 1.602e-19, 1.660538921e-27
QM = QE / (16 * AMU)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def coarse_path(tmp_path_factory):
    d = tmp_path_factory.mktemp("coarse")
    p = d / "mesh.dat"
    with gzip.open(os.path.join(GOLDEN, "coarse_mesh.dat.gz"), "rb") as f, open(p, "wb") as g:
        shutil.copyfileobj(f, g)
    return str(p)


def bary_all(tets, x):
    """lc of point x in every tet (reference formula, vectorised; for brute force only)."""
    return ((1.0 / 6.0) * (tets["alpha"] - x[0] * tets["beta"] + x[1] * tets["gamma"] - x[2] * tets["delta"])
            / tets["volume"][:, None])


# ---- tables ------------------------------------------------------------------------------

def test_coarse_tables_match_reference_hashes(mp, coarse_path):
    g = json.load(open(os.path.join(GOLDEN, "coarse_tables.json")))
    m = mp.Mesh.load_dat(coarse_path)
    assert (m.n_nodes, m.n_tets) == (g["n_nodes"], g["n_tets"])
    assert sha(m.tets) == g["tets_sha256"]
    assert sha(m.node_pos) == g["node_pos_sha256"]
    assert sha(m.node_volume) == g["node_volume_sha256"]
    assert sha(m.node_type) == g["node_type_sha256"]


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples absent")
@pytest.mark.parametrize("case", ["coarse", "medium"])
def test_tables_match_reference_loader(mp, orc, case):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    path = f"{REF_EXAMPLES}/one_stream/{case}/mesh.dat"
    R = orc.ref()
    assert R.ref_load_volume(path.encode()) == 0
    nn, nt = R.ref_num_nodes(), R.ref_num_tets()
    rt = np.zeros(nt, orc.TET_DTYPE)
    rp, rv, ry = np.zeros((nn, 3)), np.zeros(nn), np.zeros(nn, np.int32)
    R.ref_get_tables(rt.ctypes.data, rp.ctypes.data, rv.ctypes.data, ry.ctypes.data)
    m = mp.Mesh.load_dat(path)
    assert m.tets.tobytes() == rt.tobytes()
    assert m.node_volume.tobytes() == rv.tobytes()
    assert np.array_equal(m.node_type, ry)


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples absent")
def test_inlet_faces_match_reference(mp, orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    d = f"{REF_EXAMPLES}/one_stream/coarse"
    R = orc.ref()
    assert R.ref_load_volume(f"{d}/mesh.dat".encode()) == 0
    assert R.ref_load_surface(f"{d}/inlet.dat".encode(), 2, 0) == 0
    nf = R.ref_num_inlet_faces()
    con, cc = np.zeros((nf, 3), np.int32), np.zeros(nf, np.int32)
    area, u, v, nrm = np.zeros(nf), np.zeros((nf, 3)), np.zeros((nf, 3)), np.zeros((nf, 3))
    R.ref_get_inlet(con.ctypes.data, cc.ctypes.data, area.ctypes.data, u.ctypes.data, v.ctypes.data, nrm.ctypes.data)
    m = mp.Mesh.load_dat(f"{d}/mesh.dat")
    m.load_surface(f"{d}/inlet.dat", mp.NODE_INLET)
    f = m.inlet_faces
    assert len(f) == nf
    assert np.array_equal(f["con"], con) and np.array_equal(f["cell_con"], cc)
    for k, ref in (("area", area), ("u", u), ("v", v), ("normal", nrm)):
        assert f[k].tobytes() == ref.tobytes(), k


def test_box_closed_forms(mp):
    n, L = 5, 0.1
    m = mp.Mesh.box_mesh(n, L)
    t = m.tets
    h = L / n
    assert m.n_tets == 6 * n ** 3 and m.n_nodes == (n + 1) ** 3
    np.testing.assert_allclose(t["volume"], h ** 3 / 6, rtol=1e-13)
    np.testing.assert_allclose(m.node_volume.sum(), L ** 3, rtol=1e-12)
    # boundary faces: two triangles per hex face on the 6 box faces
    assert (t["cell_con"] < 0).sum() == 6 * 2 * n * n
    # neighbour table is symmetric and neighbours share exactly 3 nodes
    for e in range(0, m.n_tets, 7):
        for v in range(4):
            o = t["cell_con"][e, v]
            if o < 0:
                continue
            assert e in t["cell_con"][o]
            assert len(set(t["con"][e]) & set(t["con"][o])) == 3
            assert t["con"][e, v] not in t["con"][o]
    # lc of node j in its own tet is delta_ij
    P = m.node_pos
    for e in range(0, m.n_tets, 11):
        for j in range(4):
            lc = bary_all(t[e:e + 1], P[t["con"][e, j]])[0]
            np.testing.assert_allclose(lc, np.eye(4)[j], atol=1e-12)


def test_locate_box_brute_force(mp):
    m = mp.Mesh.box_mesh(3, 1.0)
    t = m.tets
    rng = np.random.default_rng(3)
    pts = rng.random((400, 3))
    cell = m.locate_box(pts)
    for x, c in zip(pts, cell):
        lc = bary_all(t, x)
        inside = np.flatnonzero(np.all((lc >= -1e-12) & (lc <= 1 + 1e-12), axis=1))
        assert c in inside
    out = m.locate_box(np.array([[1.5, 0.5, 0.5], [-0.1, 0.2, 0.2]]))
    assert list(out) == [-1, -1]


# ---- kernel body -----------------------------------------------------------------------

def _interior_particles(mp, m, n, seed, margin=0.3):
    """Particles at the centre region of random tets (far from faces)."""
    rng = np.random.default_rng(seed)
    t, P = m.tets, m.node_pos
    cells = rng.integers(0, m.n_tets, n).astype(np.int32)
    w = 0.25 + margin * 0.1 * (rng.random((n, 4)) - 0.5)
    w /= w.sum(1, keepdims=True)
    pos = np.einsum("nk,nkd->nd", w, P[t["con"][cells]])
    return pos, cells


def test_push_exact_arithmetic(mp, orc):
    """v' = v + (qm E) dt and x' = x + v' dt, each op correctly rounded (checked in exact rationals)."""
    m = mp.Mesh.box_mesh(2, 1.0)
    pos, cells = _interior_particles(mp, m, 64, 1)
    rng = np.random.default_rng(2)
    vel = rng.normal(size=(64, 3)) * 10.0
    ef = rng.normal(size=(m.n_tets, 3)) * 1e-3
    dt = 1e-9
    p, v, c = pos.copy(), vel.copy(), cells.copy()
    orc.move(m.tets, ef, p, v, c, QM, dt, m.n_nodes)
    assert np.array_equal(c, cells)  # no crossing at this dt

    def rnd(fr):
        return float(fr)  # Fraction -> nearest double

    for i in range(64):
        for d in range(3):
            e = ef[cells[i], d]
            a = rnd(Fraction(QM) * Fraction(e))
            b = rnd(Fraction(a) * Fraction(dt))
            vn = rnd(Fraction(vel[i, d]) + Fraction(b))
            assert v[i, d] == vn
            s = rnd(Fraction(vn) * Fraction(dt))
            assert p[i, d] == rnd(Fraction(pos[i, d]) + Fraction(s))
    # E = 0: exactly x + v dt, v unchanged
    p, v, c = pos.copy(), vel.copy(), cells.copy()
    orc.move(m.tets, np.zeros_like(ef), p, v, c, QM, dt, m.n_nodes)
    assert np.array_equal(v, vel) and np.array_equal(p, pos + vel * dt)


def test_walk_brute_force_and_loss(mp, orc):
    m = mp.Mesh.box_mesh(3, 1.0)
    t = m.tets
    rng = np.random.default_rng(5)
    n = 600
    pos = rng.random((n, 3))
    cell = m.locate_box(pos)
    vel = rng.normal(size=(n, 3)) * 0.6  # up to a few cells per step at dt = 1
    p, v, c = pos.copy(), vel.copy(), cell.copy()
    den, st, lc = orc.move(t, np.zeros((m.n_tets, 3)), p, v, c, 0.0, 1.0, m.n_nodes, lc_out=True)
    outside = np.any((p < 0) | (p > 1), axis=1)
    for i in range(n):
        if c[i] >= 0:
            lcb = bary_all(t, p[i])
            inside = np.flatnonzero(np.all((lcb >= -1e-12) & (lcb <= 1 + 1e-12), axis=1))
            assert c[i] in inside
            assert np.all(lc[i] >= 0) and np.all(lc[i] <= 1)
    # a particle is lost only if it left the box (through the face the walk took)
    assert not np.any((c < 0) & ~outside)
    assert np.all(c[outside] < 0)
    assert st["lost"] == int((c < 0).sum()) and st["particle_steps"] == n
    # raw density sums to the number of particles still inside
    np.testing.assert_allclose(den.sum(), (c >= 0).sum(), rtol=1e-13)


def test_scatter_single_particle(mp, orc):
    m = mp.Mesh.box_mesh(2, 1.0)
    t, P = m.tets, m.node_pos
    e = 17
    x = P[t["con"][e]].mean(0)[None, :].copy()
    v = np.zeros((1, 3))
    c = np.array([e], np.int32)
    den, _, lc = orc.move(t, np.zeros((m.n_tets, 3)), x, v, c, 0.0, 1.0, m.n_nodes, lc_out=True)
    assert c[0] == e
    exp = np.zeros(m.n_nodes)
    exp[t["con"][e]] = 0.25
    np.testing.assert_allclose(den, exp, atol=1e-15)
    spwt = 200.0
    out = orc.normalise(den, m.node_volume, spwt)
    np.testing.assert_allclose(out, den * spwt / m.node_volume, rtol=1e-15)


def test_compaction_stable(orc):
    n = 50
    rng = np.random.default_rng(9)
    pos, vel = rng.random((n, 3)), rng.random((n, 3))
    cell = rng.integers(-1, 5, n).astype(np.int32)
    keep = cell >= 0
    p2, v2, c2, pid = orc.compact(pos.copy(), vel.copy(), cell.copy())
    assert np.array_equal(pid, np.flatnonzero(keep))
    assert np.array_equal(p2, pos[keep]) and np.array_equal(c2, cell[keep])


def test_walk_cap(mp, orc):
    m = mp.Mesh.box_mesh(4, 1.0)
    rng = np.random.default_rng(11)
    pos = rng.random((300, 3)) * 0.5 + 0.25
    cell = m.locate_box(pos)
    vel = rng.normal(size=(300, 3)) * 0.3
    p, v, c = pos.copy(), vel.copy(), cell.copy()
    _, st, _ = orc.move(m.tets, np.zeros((m.n_tets, 3)), p, v, c, 0.0, 1.0, m.n_nodes, walk_cap=1)
    p2, v2, c2 = pos.copy(), vel.copy(), cell.copy()
    _, st2, _ = orc.move(m.tets, np.zeros((m.n_tets, 3)), p2, v2, c2, 0.0, 1.0, m.n_nodes, walk_cap=4096)
    # capped particles are exactly those whose uncapped walk needed more than one move
    assert st["cap_hits"] > 0
    assert st["cap_hits"] + st2["lost"] - st["lost"] == 0 or st["lost"] >= st["cap_hits"]
    assert np.all(c[c2 < 0] < 0)


def test_reflect_closed_form(mp, orc):
    L = 1.0
    m = mp.Mesh.box_mesh(4, L)
    x = np.array([[0.95, 0.4, 0.6]])
    v = np.array([[0.1, 0.0, 0.0]])  # crosses x = 1 by 0.05 in one unit step
    c = m.locate_box(x)
    p, vv, cc = x.copy(), v.copy(), c.copy()
    _, st, _ = orc.move(m.tets, np.zeros((m.n_tets, 3)), p, vv, cc, 0.0, 1.0, m.n_nodes, wall_mode=1)
    assert cc[0] >= 0 and st["reflections"] == 1
    np.testing.assert_allclose(p[0], [2 * L - 1.05, 0.4, 0.6], atol=1e-14)
    np.testing.assert_allclose(vv[0], [-0.1, 0.0, 0.0], atol=1e-15)
    # many particles: reflecting walls lose nothing and keep everyone in the box, |v| kept
    rng = np.random.default_rng(4)
    pos = rng.random((500, 3))
    vel = rng.normal(size=(500, 3)) * 0.2
    cell = m.locate_box(pos)
    p, vv, cc = pos.copy(), vel.copy(), cell.copy()
    for _ in range(5):
        _, st, _ = orc.move(m.tets, np.zeros((m.n_tets, 3)), p, vv, cc, 0.0, 1.0, m.n_nodes, wall_mode=1)
        assert st["lost"] == 0
    assert np.all((p >= -1e-12) & (p <= L + 1e-12))
    np.testing.assert_allclose(np.linalg.norm(vv, axis=1), np.linalg.norm(vel, axis=1), rtol=1e-12)


# ---- reference code and golden vectors ---------------------------------------------

def test_oracle_matches_golden_coarse(mp, orc, coarse_path):
    g = np.load(os.path.join(GOLDEN, "coarse_move.npz"))
    m = mp.Mesh.load_dat(coarse_path)
    parts = np.zeros(len(g["cell0"]), orc.PARTICLE_DTYPE)
    parts["pos"], parts["vel"], parts["cell_index"] = g["pos0"], g["vel0"], g["cell0"]
    for k in range(3):
        den, _ = orc.move_aos(m.tets, g["ef"], parts, float(g["qm"]), float(g["dt"]), np.zeros(m.n_nodes))
        assert np.array_equal(np.flatnonzero(parts["cell_index"] < 0), g[f"lost{k}"])
        parts = orc.compact_aos(parts).copy()
        assert parts.tobytes() == g[f"parts{k}"].tobytes()
        assert orc.normalise(den, m.node_volume, float(g["spwt"])).tobytes() == g[f"den{k}"].tobytes()


def test_oracle_matches_reference_xtoltet_box(mp, orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    m = mp.Mesh.box_mesh(6, 0.2)
    rng = np.random.default_rng(12)
    n = 5000
    pos = rng.random((n, 3)) * 0.2
    parts = np.zeros(n, orc.PARTICLE_DTYPE)
    parts["pos"], parts["vel"], parts["cell_index"] = pos, rng.normal(size=(n, 3)) * 4e4, m.locate_box(pos)
    ef = m.field_sinusoid(10.0) * 100
    tets = m.tets  # keep the table alive while the harness reads it
    a, b = parts.copy(), parts.copy()
    da, _ = orc.move_aos(tets, ef, a, QM, 1e-7, np.zeros(m.n_nodes))
    db = np.zeros(m.n_nodes)
    orc.ref().ref_move_serial(tets.ctypes.data, ef.ctypes.data, b.ctypes.data, n, QM, 1e-7, db.ctypes.data)
    assert a.tobytes() == b.tobytes()
    assert da.tobytes() == db.tobytes()


def test_rnd_golden(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built")
    g = np.load(os.path.join(GOLDEN, "rnd_mt19937.npz"))["draws"]
    out = np.zeros(64)
    orc.ref().ref_rnd(64, out.ctypes.data)
    assert np.array_equal(out, g)
    # std::mt19937(0) + uniform_real_distribution<double>: 2 x 32-bit words per draw
    assert 0.0 <= g.min() and g.max() < 1.0


# ---- adjoint oracle --------------------------------------------------------------------

@pytest.mark.parametrize("wall", [0, 1], ids=["absorb", "reflect"])
def test_adjoint_oracle_vs_finite_differences(mp, orc, wall):
    """The oracle's gradients against central differences of the serial forward run, with
    absorbing walls and with reflecting ones (the particles are fast enough that many of them
    reflect, several more than once; the perturbations keep every path and reflection)."""
    m = mp.Mesh.box_mesh(4, 1.0)
    rng = np.random.default_rng(21)
    n = 400
    pos = rng.random((n, 3)) * 0.8 + 0.1
    cell = m.locate_box(pos)
    drift = np.array([1.0, 0.0, 0.0])
    u = rng.normal(size=(n, 3)) * (0.02 if wall == 0 else 0.2)
    ef = m.field_sinusoid(1.0)
    qm, dt, spwt, K = 0.05, 0.5, 1.0, 6
    tgt = rng.random(m.n_nodes) * 5

    def fwd(e, dr):
        run = orc.Run(m.tets, e, m.node_volume, pos, u + dr * drift, cell, wall_mode=wall)
        for _ in range(K):
            den = run.step(qm, dt, spwt)
        fwd.refl = run.stats["reflections"]
        return float(((den - tgt) ** 2).sum()), run.cell.copy(), run.pid.copy()

    L0, g, gd = orc.adjoint(m.tets, ef, m.node_volume, pos, u + 0.1 * drift, cell, qm, dt, spwt, K, tgt, drift,
                            wall_mode=wall)
    Lf, c0, pid0 = fwd(ef, 0.1)
    assert abs(L0 - Lf) <= 1e-12 * Lf
    if wall:
        assert fwd.refl > n // 4 and len(c0) == n  # many reflections, nobody lost
    h = 1e-6  # rounding error eps * L / h stays far below the tolerance at L ~ 1e7
    idx = rng.choice(np.flatnonzero(np.abs(g).sum(1) > 0), 12, replace=False)
    for e in idx:
        for d in range(3):
            ep, em = ef.copy(), ef.copy()
            ep[e, d] += h
            em[e, d] -= h
            Lp, cp, pp = fwd(ep, 0.1)
            Lm, cm, pm = fwd(em, 0.1)
            assert np.array_equal(cp, c0) and np.array_equal(cm, c0)  # same cell paths
            fd = (Lp - Lm) / (2 * h)
            assert abs(fd - g[e, d]) <= 1e-5 * max(1.0, abs(g[e, d])), (e, d, fd, g[e, d])
    hp = 1e-6
    fd = (fwd(ef, 0.1 + hp)[0] - fwd(ef, 0.1 - hp)[0]) / (2 * hp)
    assert abs(fd - gd) <= 1e-5 * max(1.0, abs(gd))
