"""GPU checks of the host entry points over the C ABI: partial periodicity
(zero lattice rows), the Python calculator and the calc-efs / calc-errors
command line, each against the FP64 oracle."""
import io

import numpy as np
import pytest

from conftest import close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_available):
    from superlinear_amd import _lib
    _lib.lib()
    return _lib


@pytest.mark.parametrize("axes", [1, 2])
def test_partial_periodicity_matches_oracle(lib, oracle_lib, axes):
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print(f"[case] partial periodicity axes={axes}", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = (W.partial_frames(2, 8, 20, 2, seed=10 + axes, axes=axes) + W.periodic_frames(1, 8, 12, 2, 0.1, seed=3)
              + W.molecule_frames(1, 8, 12, 2, seed=4))
    ref = oracle_lib.OracleModel(path).efs(frames, nthreads=8)
    r = Sus2Model(path).calc_efs(frames, "fp64")
    close(r["energy"], ref["energy"], 1e-10)
    close(r["forces"], ref["forces"], 1e-10)
    close(r["virial"], ref["virial"], 1e-10)


def test_calculator_matches_oracle(lib, oracle_lib):
    from superlinear_amd import workloads as W
    from superlinear_amd.calculator import SuperlinearModel
    print("[case] calculator", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    m = SuperlinearModel(path, [6, 14])
    om = oracle_lib.OracleModel(path)
    f = W.periodic_frames(1, 10, 14, 2, 0.1, seed=21)[0]
    numbers = np.where(f["types"] == 0, 6, 14)
    out = m.calculate(numbers, f["positions"], f["cell"], [True, True, True])
    ref = om.efs([f])
    close(out["energy"], ref["energy"][0], 1e-10)
    close(out["forces"], ref["forces"], 1e-10)
    w = ref["virial"][0]
    V = np.array([[w[0], w[3], w[4]], [w[3], w[1], w[5]], [w[4], w[5], w[2]]])
    close(out["virial"], V, 1e-10)
    vol = abs(np.linalg.det(f["cell"]))
    close(out["volume"], vol, 1e-12)
    close(out["stress"], -np.array([V[0, 0], V[1, 1], V[2, 2], V[1, 2], V[0, 2], V[0, 1]]) / vol, 1e-10)
    # slab: the third row is not periodic (zeroed), the volume still uses the given cell
    out2 = m.calculate(numbers, f["positions"], f["cell"], [True, True, False])
    g = dict(f)
    g["cell"] = f["cell"] * np.array([[1], [1], [0]])
    ref2 = om.efs([g])
    close(out2["energy"], ref2["energy"][0], 1e-10)
    close(out2["forces"], ref2["forces"], 1e-10)
    # molecule: no stress for a zero cell
    out3 = m.calculate(numbers, f["positions"], np.zeros((3, 3)), [False] * 3)
    assert out3["stress"] is None


def test_cli_calc_efs_and_errors(lib, oracle_lib, tmp_path):
    from superlinear_amd import workloads as W
    from superlinear_amd.cfg import read_cfg, write_cfg
    from superlinear_amd.cli import calc_efs, calc_errors, main
    print("[case] cli", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = W.periodic_frames(3, 6, 12, 2, 0.1, seed=31) + W.molecule_frames(1, 6, 9, 2, seed=32) + \
        W.partial_frames(1, 6, 10, 2, seed=33, axes=2)
    inp = tmp_path / "in.cfg"
    with open(inp, "w") as fh:
        for f in frames:
            write_cfg(fh, f, lossless=True)
    out = tmp_path / "out.cfg"
    assert main(["calc-efs", str(path), str(inp), str(out), "--no-loss", "--batch-atoms", "20"]) == 0
    om = oracle_lib.OracleModel(path)
    ref = om.efs(frames)
    got = list(read_cfg(out))
    assert len(got) == len(frames)
    a = 0
    off = om.offsets()
    th = om.coeffs()
    for k, g in enumerate(got):
        n = len(g["types"])
        close(g["energy"], ref["energy"][k], 1e-10)
        close(g["forces"], ref["forces"][a:a + n], 1e-10)
        close(g["virial"], ref["virial"][k], 1e-10)
        const = th[off["species"] + g["types"]] + th[off["shift"] + g["types"]]
        close(g["site_energy"], ref["site"][a:a + n] - const, 1e-10)
        a += n
    # errors against the oracle's own labels vanish; against shifted labels they are known
    db = tmp_path / "db.cfg"
    with open(db, "w") as fh:
        a = 0
        for k, f in enumerate(frames):
            n = len(f["types"])
            g = dict(f, energy=ref["energy"][k] + 0.01 * n, forces=ref["forces"][a:a + n] + 0.002,
                     virial=ref["virial"][k] + 0.5)
            write_cfg(fh, g, lossless=True)
            a += n
    buf = io.StringIO()
    m = calc_errors(str(path), str(db), stream=buf)
    assert "Energy MAE (meV/atom)" in buf.getvalue()
    close(m["energy_mae_mev_atom"], 10.0, 1e-6)
    close(m["energy_rmse_mev_atom"], 1000 * np.sqrt(sum(0.01 ** 2 * len(f["types"]) for f in frames)
                                                   / sum(len(f["types"]) for f in frames)), 1e-6)
    close(m["force_mae_mev_a"], 2.0, 1e-6)
    close(m["stress_rmse_ev"], 0.5, 1e-9)
    assert calc_efs(str(path), str(inp), str(tmp_path / "o32.cfg"), precision="fp32") == len(frames)


def test_cpp_cli_matches_python_cli(lib, tmp_path):
    """The C++ driver (superlinear_amd/mlp-sus2, over the C ABI) writes the
    same lossless configurations as the Python calc-efs on the same batches,
    and prints the same calc-errors summary block."""
    import subprocess
    from superlinear_amd import workloads as W
    from superlinear_amd.build import CLI
    from superlinear_amd.cfg import read_cfg, write_cfg
    from superlinear_amd.cli import calc_efs, calc_errors
    print("[case] cpp cli", flush=True)
    if not CLI.exists():
        pytest.fail("mlp-sus2 is not built (run superlinear_amd.build first)")
    path = W.model_path("effect_l3k4r2o2")
    frames = W.periodic_frames(3, 6, 12, 2, 0.1, seed=71) + W.molecule_frames(1, 6, 9, 2, seed=72) + \
        W.partial_frames(1, 6, 10, 2, seed=73, axes=2)
    inp = tmp_path / "in.cfg"
    with open(inp, "w") as fh:
        for k, f in enumerate(frames):
            n = len(f["types"])
            g = dict(f, energy=-3.0 * n + 0.1 * k, forces=np.full((n, 3), 0.01 * (k + 1)),
                     virial=np.linspace(-1.0, 1.0, 6) * (k + 1))
            write_cfg(fh, g, lossless=True)
    out_py, out_cc = tmp_path / "py.cfg", tmp_path / "cc.cfg"
    calc_efs(str(path), str(inp), str(out_py), lossless=True, max_atoms=20)
    r = subprocess.run([str(CLI), "calc-efs", str(path), str(inp), str(out_cc), "--no-loss", "--batch-atoms", "20"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, b = list(read_cfg(out_py)), list(read_cfg(out_cc))
    assert len(a) == len(b) == len(frames)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x["types"], y["types"])
        for k in ("energy", "forces", "virial", "site_energy", "positions"):
            close(y[k], x[k], 1e-14)
    buf = io.StringIO()
    calc_errors(str(path), str(inp), max_atoms=20, stream=buf)
    r = subprocess.run([str(CLI), "calc-errors", str(path), str(inp), "--batch-atoms", "20"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == buf.getvalue().strip()
