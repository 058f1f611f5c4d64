"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run on the build host, where /root/reference exists and oracle/_ref/libref.so has been
built from the reference sources (make -C oracle ref).  The reference ships no tests or
golden vectors, so these fixtures are produced by its compiled src/meshes.cpp
(LoadVolumeMesh, XtoLtet) and src/maths.cpp (det3/det4, rnd) through oracle/ref_harness.cpp:

  coarse_mesh.dat.gz   examples/one_stream/coarse/mesh.dat (input data, gzipped)
  coarse_tables.json   SHA-256 of LoadVolumeMesh's Tetra table, node volumes and node types
  coarse_move.npz      3 MoveParticles steps on the coarse cylinder: seeded inputs, the
                       reference kernel body (its XtoLtet) + stable erase + normalisation
  rnd_mt19937.npz      first draws of the reference's rnd()
"""
import gzip
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

REF = "/root/reference/examples/one_stream/coarse"
QE, AMU = 1.602e-19, 1.660538921e-27


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_tables(path):
    R = oracle.ref()
    assert R.ref_load_volume(path.encode()) == 0
    nn, nt = R.ref_num_nodes(), R.ref_num_tets()
    tets = np.zeros(nt, oracle.TET_DTYPE)
    pos = np.zeros((nn, 3))
    vol = np.zeros(nn)
    typ = np.zeros(nn, np.int32)
    R.ref_get_tables(tets.ctypes.data, pos.ctypes.data, vol.ctypes.data, typ.ctypes.data)
    return tets, pos, vol, typ


def main():
    R = oracle.ref()
    src = os.path.join(REF, "mesh.dat")
    with open(src, "rb") as f, gzip.open(os.path.join(HERE, "coarse_mesh.dat.gz"), "wb", compresslevel=9) as g:
        shutil.copyfileobj(f, g)
    tets, pos, vol, typ = ref_tables(src)
    json.dump({"source": "examples/one_stream/coarse/mesh.dat via src/meshes.cpp:28-265",
               "n_nodes": len(vol), "n_tets": len(tets), "tets_sha256": sha(tets), "node_pos_sha256": sha(pos),
               "node_volume_sha256": sha(vol), "node_type_sha256": sha(typ)},
              open(os.path.join(HERE, "coarse_tables.json"), "w"), indent=1)

    # seeded particles inside random tets, random per-element field
    rng = np.random.Generator(np.random.PCG64(20240611))
    n = 2048
    cells = rng.integers(0, len(tets), n).astype(np.int32)
    w = rng.exponential(size=(n, 4))
    w /= w.sum(1, keepdims=True)
    p0 = np.einsum("nk,nkd->nd", w, pos[tets["con"][cells]])
    v0 = rng.normal(size=(n, 3)) * 2.0e4
    ef = rng.normal(size=(len(tets), 3)) * 2.0e3
    qm, dt, spwt = QE / (16 * AMU), 1.0e-8, 200.0
    parts = np.zeros(n, oracle.PARTICLE_DTYPE)
    parts["pos"], parts["vel"], parts["cell_index"] = p0, v0, cells
    out = {"pos0": p0, "vel0": v0, "cell0": cells, "ef": ef, "qm": qm, "dt": dt, "spwt": spwt}
    cur = parts.copy()
    for k in range(3):
        den = np.zeros(len(vol))
        R.ref_move_serial(tets.ctypes.data, ef.ctypes.data, cur.ctypes.data, len(cur), qm, dt, den.ctypes.data)
        out[f"lost{k}"] = np.flatnonzero(cur["cell_index"] < 0).astype(np.int64)
        cur = cur[cur["cell_index"] >= 0].copy()          # std::erase_if, stable
        out[f"den{k}"] = den * (spwt / vol)               # src/mini-pic.cpp:319
        out[f"parts{k}"] = cur.copy()
    np.savez_compressed(os.path.join(HERE, "coarse_move.npz"), **out)

    draws = np.zeros(64)
    R.ref_rnd(64, draws.ctypes.data)
    np.savez_compressed(os.path.join(HERE, "rnd_mt19937.npz"), draws=draws)
    print("fixtures written", {k: v.shape if hasattr(v, "shape") else v for k, v in out.items()})


if __name__ == "__main__":
    main()
