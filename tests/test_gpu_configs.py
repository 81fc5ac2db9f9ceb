"""GPU checks at the benchmark workloads' own sizes, through the C ABI:
the 64-atom Static frame and the 4096-atom Effect frame against the FP64
oracle, 16 frames of the 256-frame training batch (loss and a sample of the
gradient entries against the oracle), rotation invariance of the L4 / L5
models, and size-independent properties of the 2,097,152-atom frame."""
import numpy as np
import pytest

from conftest import voigt_to_matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_available):
    from superlinear_amd import _lib
    _lib.lib()
    return _lib


def _maxrel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("config", ["static64", "effect4096"])
def test_config_frame_against_oracle(lib, oracle_lib, config):
    """FP64 within 1e-10 (relative to the largest entry) of the oracle; FP32
    forces within 1e-5 eV/A, energy within 1e-6 eV/atom, virial within
    1e-4 eV (the precision bounds of the FP32 path)."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print(f"[case] {config} frame vs oracle", flush=True)
    path = W.model_path(W.CONFIG_MODEL[config])
    frames = W.config_frames(config)
    n = len(frames[0]["types"])
    ref = oracle_lib.OracleModel(path).efs(frames, nthreads=16)
    m = Sus2Model(path)
    b = DeviceBatch(frames)
    r64 = {k: v.cpu().numpy() for k, v in m.efs_device(b, "fp64").items()}
    r32 = {k: v.cpu().numpy() for k, v in m.efs_device(b, "fp32").items()}
    e64 = abs(r64["energy"][0] - ref["energy"][0]) / abs(ref["energy"][0])
    f32 = np.abs(r32["forces"] - ref["forces"]).max()
    print(f"[result] {config}: fp64 E rel {e64:.2e} F {_maxrel(r64['forces'], ref['forces']):.2e}; "
          f"fp32 F maxabs {f32:.2e} E/atom {abs(r32['energy'][0] - ref['energy'][0]) / n:.2e}", flush=True)
    assert e64 <= 1e-10
    assert _maxrel(r64["forces"], ref["forces"]) <= 1e-10
    assert _maxrel(r64["virial"], ref["virial"]) <= 1e-10
    assert _maxrel(r64["site"], ref["site"]) <= 1e-10
    assert f32 <= 1e-5
    assert abs(r32["energy"][0] - ref["energy"][0]) / n <= 1e-6
    assert np.abs(r32["virial"] - ref["virial"]).max() <= 1e-4


def _gradient_sample(ncoeff, offsets, rng, count=40):
    """The first entry of every coefficient block plus a random sample."""
    firsts = [v for k, v in offsets.items() if k != "ncoeff" and v >= 0]
    return np.unique(np.concatenate([firsts, rng.choice(ncoeff, count, replace=False)])).astype(np.int32)


def test_training_frames_against_oracle(lib, oracle_lib):
    """16 frames of the training batch (every 16th of the 256): the objective
    (weighting structures, wE 1, wF 0.01, wS 0) in FP64 within 1e-12 and FP32
    within 1e-6 of the oracle; gradient entries of every coefficient block
    and a random sample: FP64 within 1e-10, the mixed objective within 1e-6,
    FP32 within 1e-5 (relative, vector norm over the sample)."""
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print("[case] train256 sample vs oracle", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = W.config_frames("train256")[::16]
    assert len(frames) == 16
    om = oracle_lib.OracleModel(path)
    theta = om.coeffs()
    rng = np.random.default_rng(3031)
    om.set_coeffs(theta + 0.05 * np.abs(theta) * rng.standard_normal(theta.size))
    teach = om.efs(frames, nthreads=16)
    om.set_coeffs(theta)
    labels = dict(energy=teach["energy"], forces=teach["forces"])
    idx = _gradient_sample(theta.size, om.offsets(), np.random.default_rng(7))
    om.grad_subset(idx)
    ref = om.loss_grad(frames, labels, mode=1, wE=1.0, wF=0.01, wS=0.0, nthreads=16)
    om.grad_subset(None)
    m = Sus2Model(path)
    b = DeviceBatch(frames)
    e0, f0 = torch.from_numpy(teach["energy"]).cuda(), torch.from_numpy(teach["forces"]).cuda()
    rg = ref["grad"][idx]
    for prec, ltol, gtol in (("fp64", 1e-12, 1e-10), ("mixed", 1e-12, 1e-6), ("fp32", 1e-6, 1e-5)):
        loss, g = m.loss_grad_device(b, e0, f0, weighting="structures", wE=1.0, wF=0.01, wS=0.0, precision=prec)
        g = g.cpu().numpy()[idx]
        lrel = abs(loss - ref["loss"]) / abs(ref["loss"])
        grel = float(np.linalg.norm(g - rg) / np.linalg.norm(rg))
        print(f"[result] train256 sample {prec}: loss rel {lrel:.2e} grad rel {grel:.2e} ({idx.size} entries)",
              flush=True)
        assert lrel <= ltol
        assert grel <= gtol


@pytest.mark.parametrize("mname,species", [("effect_l4k6r2o3", 3), ("effect_l5k6r2o3", 2)])
def test_rotation_invariance_high_l(lib, mname, species):
    """GPU FP64 EFS of a rotated periodic frame: energy and site energies
    equal, forces and virial rotated (L = 4, 5 couplings and harmonics)."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print(f"[case] rotation {mname}", flush=True)
    m = Sus2Model(W.model_path(mname))
    frame = W.periodic_frames(1, 60, 90, species, 0.15, seed=29)[0]
    q, _ = np.linalg.qr(np.random.default_rng(5).normal(size=(3, 3)))
    rot = dict(frame, positions=frame["positions"] @ q.T, cell=frame["cell"] @ q.T)
    for prec, tol in (("fp64", 1e-11), ("fp32", 2e-5)):
        a = m.calc_efs([frame], prec)
        b = m.calc_efs([rot], prec)
        assert _maxrel(b["energy"], a["energy"]) <= tol
        assert _maxrel(b["site"], a["site"]) <= tol
        assert _maxrel(b["forces"], a["forces"] @ q.T) <= tol
        assert _maxrel(voigt_to_matrix(b["virial"][0]), q @ voigt_to_matrix(a["virial"][0]) @ q.T) <= tol


def test_md2m_frame_properties(lib):
    """The 2,097,152-atom frame (device-wide binning path, 1.8e8 edges):
    undisplaced lattice has exactly 86 neighbours per atom (shells
    4 + 12 + 12 + 6 + 12 + 24 + 16 within 5 A); on the displaced frame
    sum F = 0, E = sum E_i, FP32 forces within 1e-5 eV/A of FP64, and a
    rigid translation leaves energy and forces unchanged."""
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print("[case] md2m properties", flush=True)
    m = Sus2Model(W.model_path("effect_l5k6r2o3"))
    lattice = W.diamond(64, sigma=0.0, species=2, seed=505)
    n = len(lattice["types"])
    assert n == 2097152
    m.efs_device(DeviceBatch([lattice]), "fp32")
    torch.cuda.synchronize()
    assert m.last_edge_count() == 86 * n
    frame = W.config_frames("md2m")[0]
    b = DeviceBatch([frame])
    r32 = {k: v.cpu().numpy() for k, v in m.efs_device(b, "fp32").items()}
    edges = m.last_edge_count()
    assert edges % 2 == 0
    r64 = {k: v.cpu().numpy() for k, v in m.efs_device(b, "fp64").items()}
    assert m.last_edge_count() == edges
    shifted = dict(frame, positions=frame["positions"] + np.array([0.37, -1.21, 2.9]))
    r64s = {k: v.cpu().numpy() for k, v in m.efs_device(DeviceBatch([shifted]), "fp64").items()}
    fmax = np.abs(r64["forces"]).max()
    df = np.abs(r32["forces"] - r64["forces"]).max()
    print(f"[result] md2m: {edges} edges, fp32-fp64 F maxabs {df:.2e}, |sum F| {np.abs(r64['forces'].sum(0)).max():.2e}",
          flush=True)
    assert np.abs(r64["forces"].sum(0)).max() <= 1e-9 * fmax * np.sqrt(n)
    assert abs(r64["site"].sum() - r64["energy"][0]) <= 1e-12 * np.abs(r64["site"]).sum()
    assert df <= 1e-5
    assert abs(r32["energy"][0] - r64["energy"][0]) / n <= 1e-6
    assert abs(r64s["energy"][0] - r64["energy"][0]) <= 1e-10 * abs(r64["energy"][0])
    assert np.abs(r64s["forces"] - r64["forces"]).max() <= 1e-10 * fmax
