import json
import os
import re
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"
FIXTURES = GOLDEN / "two_cutoff"
os.environ.setdefault("SUS2_MODEL_DIR", str(ROOT / "superlinear_amd" / "_models"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsus2amd.so")
    config.addinivalue_line("markers", "slow: longer CPU checks")


def voigt_to_matrix(w):
    return np.array([[w[0], w[3], w[4]], [w[3], w[1], w[5]], [w[4], w[5], w[2]]])


def fixture_cases():
    """(name, model path, frame, expected) for every record of the bundled
    reference FP64 fixtures (interfaces/lammps/tests/fixtures/two_cutoff)."""
    from superlinear_amd.cfg import read_cfg
    out = []
    for d, names in (("fixtures_v2", ("static", "effect")), ("fixtures_hi_l", ("effect_l3k4r3o3",))):
        geos = list(read_cfg(FIXTURES / d / "geometries.cfg"))
        for nm in names:
            ref = json.loads((FIXTURES / d / f"{nm}.efs.json").read_text())
            for rec in ref["records"]:
                out.append((f"{nm}/frame{rec['config']}", FIXTURES / d / f"{nm}.mtp", geos[rec["config"]], rec))
    return out


def close(actual, expected, tol):
    a = np.asarray(actual, float).ravel()
    b = np.asarray(expected, float).ravel()
    assert a.shape == b.shape
    err = np.abs(a - b)
    lim = tol * (1 + np.maximum(np.abs(a), np.abs(b)))
    assert np.all(err <= lim), f"max abs err {err.max():.3e}"
    return float(err.max(initial=0))


def real_sh_scipy(L, d):
    """Real Y_lm (l <= L, flat index l^2 + l + m) from scipy's complex
    harmonics: m > 0 cos, m < 0 sin parts with the Condon-Shortley phase
    cancelled, so that Y_1,(-1, 0, 1) is proportional to (y, z, x)."""
    from scipy.special import sph_harm_y
    r = np.linalg.norm(d)
    theta, phi = np.arccos(d[2] / r), np.arctan2(d[1], d[0])
    out = np.zeros((L + 1) ** 2)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)  # includes the Condon-Shortley phase
            cs = (-1) ** abs(m)
            if m == 0:
                v = y.real
            elif m > 0:
                v = np.sqrt(2) * cs * y.real
            else:
                v = np.sqrt(2) * cs * y.imag
            out[l * l + l + m] = v
    return out


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    try:
        oracle.lib()
    except RuntimeError:
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def gpu_available():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return True
