"""Batched assembly: every mesh of a batch gets exactly the CSR it gets when
assembled alone (pattern and values bit-identical), which the oracle checks as
usual; graph replays are deterministic; the autotuner keeps that true."""

import numpy as np
import pytest

from finite_element_method_amd import FEM, FEMBatch, meshgen
from oracle.oracle import OracleModel
from tests.parity import compare, oracle_reference

pytestmark = pytest.mark.gpu


def _tilt():
    a = 0.4
    return np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])


def _meshes():
    return [meshgen.plate_grid(6, seed=1), meshgen.truss_lattice(3, seed=2), meshgen.plate_grid(5, 4, seed=3, shuffle=True),
            meshgen.beam_frame(3, seed=4), meshgen.plate_grid(4, seed=5, rotation=_tilt()),
            meshgen.plate_grid(9, 3, seed=6), meshgen.truss_lattice(4, seed=7), meshgen.beam_frame(2, 3, 2, seed=8)]


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_batch_matches_single_assembly_and_oracle(gpu):
    meshes = _meshes()
    models = [FEM.from_mesh(m) for m in meshes]
    batch = FEMBatch(models)
    assert batch.groups() == 4  # flat plates | bars with beams | truss-only meshes | the rest
    batch.assemble()
    for i, (mesh, model) in enumerate(zip(meshes, models)):
        print("mesh", i, mesh.kind, mesh.n_elements, flush=True)
        got = batch.mesh_csr(i)
        alone = model.stiffness_csr()
        assert _same(got, alone), i
        r = compare(got, oracle_reference(meshgen.load_oracle(mesh, OracleModel)))
        assert r["pattern_equal"] and r["max_ratio"] <= 1.0, (i, r)


def test_batch_autotune_and_replay(gpu):
    meshes = [meshgen.plate_grid(n, seed=n) for n in (3, 17, 8)] + [meshgen.beam_frame(4, seed=9),
                                                                   meshgen.truss_lattice(5, seed=10)]
    models = [FEM.from_mesh(m) for m in meshes]
    batch = FEMBatch(models)
    cfg = batch.autotune()
    print("config", cfg, flush=True)
    assert "plate_lanes" in cfg and "bar_owner" in cfg
    batch.assemble()
    first = [batch.mesh_csr(i) for i in range(len(meshes))]
    for _ in range(3):
        batch.assemble()
    assert batch.last_ms() > 0
    for i, model in enumerate(models):
        again = batch.mesh_csr(i)
        assert _same(again, first[i])
        assert _same(again, model.stiffness_csr())
    # a second batch with the same signature reuses the tuned shapes
    batch2 = FEMBatch(models)
    assert batch2.autotune() == cfg


_CACHE_SCRIPT = r"""
import sys
sys.path.insert(0, %r)
from finite_element_method_amd import FEM, FEMBatch, meshgen
meshes = [meshgen.plate_grid(5, seed=31), meshgen.truss_lattice(3, seed=32), meshgen.beam_frame(3, seed=33)]
batch = FEMBatch([FEM.from_mesh(m) for m in meshes])
print("CONFIG", batch.autotune(), flush=True)
"""


def test_tune_cache_persists(gpu, tmp_path):
    """The tuned shapes and schedule are stored per (gfx arch, batch signature)
    in the tune cache file, and a later process applies a stored entry
    without timing: flipping the stored schedule flips the next process's."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, FEM_AMD_TUNE_CACHE=str(cache))
    script = _CACHE_SCRIPT % root

    def run():
        out = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
        return [ln for ln in out.stdout.splitlines() if ln.startswith("CONFIG")][0]

    first = run()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1
    arch, sig, sched, cfgs = lines[0].split()
    assert arch.startswith("gfx") and sched in ("0", "1") and len(cfgs.split(";")) == 3
    assert ("groups serial" if sched == "1" else "groups concurrent") in first
    flipped = "0" if sched == "1" else "1"
    cache.write_text(" ".join([arch, sig, flipped, cfgs]) + "\n")
    second = run()
    assert ("groups serial" if flipped == "1" else "groups concurrent") in second
    assert cache.read_text().splitlines() == [" ".join([arch, sig, flipped, cfgs])]  # nothing re-tuned
