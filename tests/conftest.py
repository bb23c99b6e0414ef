import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def potentials(tmp_path_factory):
    """ann files of the workloads, written once per session."""
    from atomnn_amd import synth

    d = tmp_path_factory.mktemp("pot")
    out = {}
    for name in ("plain64", "plain4096", "mixed64x4096"):
        p = str(d / f"{name}.ann")
        synth.workload_potential(name).write(p)
        out[name] = p
    # a potential outside the moment expansion's domain (non-integer zeta): triplet loops only
    import numpy as np
    frac = synth.make_potential(["A", "B"], 6.0, 6, [0.005, 0.05], [1.5, 4.0], [1, -1], [16, 16],
                                np.random.default_rng(77))
    p = str(d / "frac_zeta.ann")
    frac.write(p)
    out["frac_zeta"] = p
    out["batch512"] = out["plain4096"]
    out["serve20k"] = out["mixed64x4096"]
    return out
