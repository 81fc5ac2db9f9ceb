"""The allocation-free EFS entry point (sus2_efs_noalloc, scratch carved from
one static device array per GPU) against sus2_efs through the C ABI: two
calls in a row, the second on a smaller batch, in both precisions, and a
batch too large for the scratch array rejected with SUS2_E_RANGE."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_available):
    from superlinear_amd import _lib
    return _lib


def _noalloc(lib, model, batch, precision):
    import torch
    from superlinear_amd.model import PRECISION
    nf, n = batch.nframes, batch.natoms
    out = dict(energy=torch.full((nf,), np.nan, dtype=torch.float64, device="cuda"),
               forces=torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda"),
               virial=torch.full((nf, 6), np.nan, dtype=torch.float64, device="cuda"),
               site=torch.full((n,), np.nan, dtype=torch.float64, device="cuda"))
    st = torch.cuda.current_stream().cuda_stream
    rc = lib.lib().sus2_efs_noalloc(model.handle, PRECISION[precision], nf, n, batch.off.data_ptr(),
                                     batch.types.data_ptr(), batch.pos.data_ptr(), batch.cell.data_ptr(),
                                     batch.periodic.data_ptr(), out["energy"].data_ptr(), out["forces"].data_ptr(),
                                     out["virial"].data_ptr(), out["site"].data_ptr(), ctypes.c_void_p(st))
    return rc, out


def _close(a, b, tol):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.all(np.isfinite(a))
    assert np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6)])
def test_noalloc_matches_efs(lib, precision, tol):
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print(f"[case] noalloc {precision}", flush=True)
    m = Sus2Model(W.model_path("effect_l3k4r2o2"))
    big = DeviceBatch(W.config_frames("effect4096"))
    small = DeviceBatch(W.periodic_frames(3, 20, 60, 2, 0.15, seed=71) + W.molecule_frames(2, 10, 30, 2, seed=72))
    for batch in (big, small):  # the second call on a smaller batch
        ref = m.efs_device(batch, precision)
        torch.cuda.synchronize()
        rc, out = _noalloc(lib, m, batch, precision)
        assert rc == 0, lib.lib().sus2_last_error().decode()
        torch.cuda.synchronize()
        for k in ("energy", "forces", "virial", "site"):
            _close(out[k], ref[k], tol)


def test_noalloc_rejects_oversized_batch_and_recovers(lib):
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print("[case] noalloc oversized", flush=True)
    m = Sus2Model(W.model_path("effect_l3k4r2o2"))
    small = DeviceBatch([W.diamond(2, species=2, seed=3)])
    ref = m.efs_device(small, "fp64")
    big = DeviceBatch([W.diamond(14, species=2, seed=4)])  # 21952 atoms: beyond the scratch array in FP64
    rc, _ = _noalloc(lib, m, big, "fp64")
    assert rc == 5, lib.lib().sus2_last_error().decode()  # SUS2_E_RANGE
    assert "static region" in lib.lib().sus2_last_error().decode()
    torch.cuda.synchronize()
    rc, out = _noalloc(lib, m, small, "fp64")
    assert rc == 0
    torch.cuda.synchronize()
    _close(out["forces"], ref["forces"], 1e-12)


def test_noalloc_needs_device_copy(lib):
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    m = Sus2Model(W.model_path("effect_l3k4r2o2"))
    rc, _ = _noalloc(lib, m, DeviceBatch([W.diamond(2, species=2, seed=3)]), "fp64")
    assert rc == 1  # SUS2_E_INVALID: no device copy yet
