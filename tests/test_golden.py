"""Stored oracle fixtures (tests/golden): the oracle reproduces them bit for bit, and the GPU
FP64 path matches them to 1e-12 relative L-inf per field."""
import hashlib
import json
import os

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi
from mrlbm_3d_cuda_amd._abi import CaseDesc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))


def desc_from(d):
    out = CaseDesc()
    for k, v in d.items():
        if isinstance(v, list):
            getattr(out, k)[:] = v
        else:
            setattr(out, k, v)
    return out


def load(name):
    z = np.load(os.path.join(HERE, name + ".npz"))
    return z["fields"], z["forces"]


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_fixture_checksums(name):
    fields, forces = load(name)
    assert hashlib.sha256(fields.tobytes()).hexdigest() == MANIFEST[name]["fields_sha256"]
    assert hashlib.sha256(forces.tobytes()).hexdigest() == MANIFEST[name]["forces_sha256"]


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_oracle_reproduces_fixture(name):
    from oracle.binding import Oracle
    m = MANIFEST[name]
    o = Oracle(desc_from(m["desc"])).init()
    o.step(m["steps"], _abi.STEP_FORCES)
    fields, forces = load(name)
    assert np.array_equal(o.fields(), fields)
    assert np.array_equal(o.forces(0, m["steps"]), forces)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_gpu_matches_fixture(name):
    from mrlbm_3d_cuda_amd.solver import Simulation
    m = MANIFEST[name]
    print(f"[gpu golden] {name}", flush=True)
    s = Simulation(desc_from(m["desc"])).init()
    s.step(m["steps"], forces=True)
    fields, forces = load(name)
    G = s.fields()
    # relative L-inf per physical field: density deviation, velocity vector, moment tensor
    for name, idx in {"rho": [0], "u": [1, 2, 3], "m": [4, 5, 6, 7, 8, 9]}.items():
        err = np.max(np.abs(G[idx] - fields[idx])) / np.max(np.abs(fields[idx]))
        print(f"[gpu golden] {name}: {err:.3e}", flush=True)
        assert err <= 1e-12, (name, err)
    F = s.forces(0, m["steps"])
    # force records against the gross momentum flux, about 2 rho0 per boundary node
    info = s.info()
    nb = info.nb_inner + info.nb_outer + info.nb_fluid_inner + info.nb_fluid_outer
    assert np.max(np.abs(F - forces)) <= 1e-12 * 2.0 * nb
