"""Regenerate the oracle fixtures in this folder (deliberately, never from a test).

    python tests/golden/make_golden.py

Each fixture holds the case descriptor, the moment fields after `steps` FP64 oracle steps from the
seeded initial state, and the per-step force records. The reference ships no golden vectors; these
pin the oracle against drift and give the GPU tests a stored expectation.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mrlbm_3d_cuda_amd import _abi, configs  # noqa: E402
from oracle.binding import Oracle  # noqa: E402

FIXTURES = {
    "annulus_24x24x4_10steps": (configs.annulus(nx=24, ny=24, nz=4, xc=12.0, yc=12.0, r_in=4.0, r_out=10.0,
                                                u_inner=0.05, tau=0.8), 10),
    "channel_32x16x4_10steps": (configs.cylinder(nx=32, ny=16, nz=4, diameter=6.0, xc=10.0, yc=8.0, u_max=0.05,
                                                 re=30.0, precision=_abi.FP64), 10),
}


def digest(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def generate(name):
    desc, steps = FIXTURES[name]
    o = Oracle(desc).init()
    o.step(steps, _abi.STEP_FORCES)
    return desc, steps, o.fields(), o.forces(0, steps)


def main():
    manifest = {}
    for name in FIXTURES:
        desc, steps, fields, forces = generate(name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), fields=fields, forces=forces)
        manifest[name] = {"desc": desc.as_dict(), "steps": steps, "fields_sha256": digest(fields),
                          "forces_sha256": digest(forces)}
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
