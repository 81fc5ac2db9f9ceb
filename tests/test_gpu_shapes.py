"""A model shape without a prebuilt kernel instance (L2/K3, radial basis 8,
scaling_map LK): its kernels come from the shape library compiled by build()
into the cache (no compiler runs here: SUS2_JIT=0), and EFS and parameter
gradients match the FP64 oracle as the prebuilt shapes do."""
import numpy as np
import pytest

from conftest import close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_available):
    from superlinear_amd import _lib
    _lib.lib()
    return _lib


def _model(tmp_path, gate):
    from superlinear_amd.topology import write_model
    path = tmp_path / ("l2k3b8lk_%s.sus2" % ("effect" if gate else "static"))
    kw = dict(gate=True, gate_rank=2, gate_pruning="q-total", outer_pruning="same-k", species=2) if gate else \
        dict(gate=False, outer_pruning="q-total", species=1)
    write_model(str(path), lmax=2, kmax=3, body_order=4, outer_rank=2, rb_size=8, scaling_map="LK", cutoff=5.0,
                seed=91 if gate else 92, **kw)
    return str(path)


@pytest.mark.parametrize("gate", [True, False], ids=["effect", "static"])
def test_runtime_shape_efs_and_grad_match_oracle(lib, oracle_lib, tmp_path, monkeypatch, gate):
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    monkeypatch.setenv("SUS2_JIT", "0")
    print(f"[case] runtime shape l2k3b8 LK gate={gate}", flush=True)
    path = _model(tmp_path, gate)
    m = Sus2Model(path)
    d = m.dims()
    assert (d["L"], d["K"], d["B"], d["Kw"]) == (2, 3, 8, 9)
    species = 2 if gate else 1
    frames = (W.periodic_frames(2, 20, 40, species, 0.12, seed=93) + W.molecule_frames(1, 8, 14, species, seed=94)
              + W.partial_frames(1, 10, 16, species, seed=95, axes=2))
    om = oracle_lib.OracleModel(path)
    ref = om.efs(frames, nthreads=8)
    r64 = m.calc_efs(frames, "fp64")
    close(r64["energy"], ref["energy"], 1e-10)
    close(r64["forces"], ref["forces"], 1e-10)
    close(r64["virial"], ref["virial"], 1e-10)
    r32 = m.calc_efs(frames, "fp32")
    assert np.abs(r32["forces"] - ref["forces"]).max() <= 1e-5
    n = sum(len(f["types"]) for f in frames)
    rng = np.random.default_rng(96)
    ew, fw, sw = rng.normal(size=n), rng.normal(size=(n, 3)), rng.normal(size=(len(frames), 9))
    gref = om.efs_grad(frames, ew, fw, sw, nthreads=8)
    batch = DeviceBatch(frames)
    g64 = m.efs_grad_device(batch, *[torch.from_numpy(x).cuda() for x in (ew, fw, sw)], precision="fp64")
    g64 = g64.cpu().numpy()
    assert np.linalg.norm(g64 - gref) <= 1e-10 * np.linalg.norm(gref)
