"""Batches captured in HIP graphs (sus2_graph_*) and the launch-shape
autotuner, through the C ABI: every replay equals the eager call on the same
inputs (with eager work and synchronisations between replays), inputs
rewritten in place between replays are picked up, a batch that outgrows the
captured neighbour capacity is reported by the device flag, and every
candidate launch shape gives the same EFS."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_available):
    from superlinear_amd import _lib
    _lib.lib()
    return _lib


def _mixed_frames(species, seed):
    from superlinear_amd import workloads as W
    return (W.periodic_frames(6, 16, 200, species, 0.10, seed=seed, log_uniform=True)
            + W.molecule_frames(6, 16, 200, species, seed=seed + 1))


def _maxrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("mname,species,prec", [("effect_l4k6r2o3", 3, "fp32"), ("effect_l4k6r2o3", 3, "fp64"),
                                                 ("static_l3k4o2", 1, "fp64")])
def test_graph_efs_replays_match_eager(lib, mname, species, prec):
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print(f"[case] graph efs {mname} {prec}", flush=True)
    m = Sus2Model(W.model_path(mname))
    batch = DeviceBatch(_mixed_frames(species, 31))
    g = m.graph_efs(batch, prec)
    rng = np.random.default_rng(5)
    base = batch.pos.clone()
    for rep in range(4):
        if rep >= 2:  # positions rewritten in place between replays
            batch.pos.copy_(base + torch.from_numpy(0.02 * rng.standard_normal(base.shape)).cuda())
        out = g.launch()
        junk = torch.randn(4096, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        ref = m.efs_device(batch, prec)
        torch.cuda.synchronize()
        assert not g.overflow()
        tol = 1e-12 if prec == "fp64" else 1e-6
        for k in ("energy", "forces", "virial", "site"):
            assert _maxrel(out[k].cpu().numpy(), ref[k].cpu().numpy()) <= tol, (rep, k)
        del junk


def test_graph_overflow_is_reported(lib):
    """Compressing the cell and positions in place by 15 % raises the
    neighbour counts beyond the capacity sized at creation: the launch empties
    the lists and sets the flag, which overflow() reports once and clears."""
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print("[case] graph overflow", flush=True)
    m = Sus2Model(W.model_path("effect_l3k4r2o2"))
    batch = DeviceBatch(W.periodic_frames(4, 60, 120, 2, 0.10, seed=37))
    g = m.graph_efs(batch, "fp32")
    g.launch()
    assert not g.overflow()
    pos, cell = batch.pos.clone(), batch.cell.clone()
    batch.pos.mul_(0.85)
    batch.cell.mul_(0.85)
    g.launch()
    assert g.overflow()
    assert not g.overflow()  # cleared
    batch.pos.copy_(pos)
    batch.cell.copy_(cell)
    out = g.launch()
    assert not g.overflow()
    ref = m.efs_device(batch, "fp32")
    torch.cuda.synchronize()
    assert _maxrel(out["forces"].cpu().numpy(), ref["forces"].cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("prec", ["fp32", "mixed", "fp64"])
def test_graph_loss_grad_replays_match_eager(lib, prec):
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print(f"[case] graph loss_grad {prec}", flush=True)
    path = W.model_path("effect_l4k6r2o3")
    batch = DeviceBatch(_mixed_frames(3, 41))
    teacher = Sus2Model(path)
    th = teacher.coeffs()
    teacher.set_coeffs(th * (1 + 0.05 * np.random.default_rng(42).standard_normal(th.size)))
    lab = teacher.efs_device(batch, "fp64")
    m = Sus2Model(path)
    g = m.graph_loss_grad(batch, lab["energy"], lab["forces"], weighting="structures", wE=1.0, wF=0.01,
                          precision=prec)
    for rep in range(3):
        if rep == 2:  # new coefficients, read by the graph
            m.set_coeffs(th * (1 + 0.01 * np.random.default_rng(43).standard_normal(th.size)))
        out = g.launch()
        torch.cuda.synchronize()
        loss, grad = m.loss_grad_device(batch, lab["energy"], lab["forces"], weighting="structures", wE=1.0,
                                        wF=0.01, precision=prec)
        torch.cuda.synchronize()
        assert not g.overflow()
        tol = 1e-12 if prec == "fp64" else 1e-9
        assert abs(out["loss"].item() - loss) <= tol * abs(loss)
        assert _maxrel(out["grad"].cpu().numpy(), grad.cpu().numpy()) <= (1e-12 if prec == "fp64" else 1e-7)
        efs = m.efs_device(batch, "fp64" if prec != "fp32" else "fp32")
        torch.cuda.synchronize()
        assert _maxrel(out["forces"].cpu().numpy(), efs["forces"].cpu().numpy()) <= 1e-6


def test_launch_shapes_agree_and_autotune(lib, oracle_lib):
    """Every candidate launch shape (centres per edge-backward group x
    persistent grid multiple) gives the oracle's EFS (FP64 1e-10) on an uneven
    batch; the autotuner times all twelve and keeps one of them."""
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print("[case] launch shapes", flush=True)
    path = W.model_path("effect_l4k6r2o3")
    frames = _mixed_frames(3, 51)
    ref = oracle_lib.OracleModel(path).efs(frames, nthreads=16)
    m = Sus2Model(path)
    batch = DeviceBatch(frames)
    for eb in (1, 2, 4):
        for gm in (0, 1, 2, 4):
            m.tune("fp64", eb, gm)
            r = {k: v.cpu().numpy() for k, v in m.efs_device(batch, "fp64").items()}
            assert _maxrel(r["forces"], ref["forces"]) <= 1e-10, (eb, gm)
            assert _maxrel(r["energy"], ref["energy"]) <= 1e-10, (eb, gm)
            assert _maxrel(r["virial"], ref["virial"]) <= 1e-10, (eb, gm)
    best, table = m.autotune(batch, "fp32", reps=3)
    print(f"[result] autotune: {best} {table}", flush=True)
    assert len(table) == 12 and all(v > 0 for v in table.values())
    assert (best["eb_atoms"], best["grid_mult"]) in table
    assert m.get_tune("fp32") == best
    r32 = {k: v.cpu().numpy() for k, v in m.efs_device(batch, "fp32").items()}
    assert np.abs(r32["forces"] - ref["forces"]).max() <= 1e-5


def test_pair_contract_with_persistent_shape(lib, oracle_lib):
    """The pair-style stages under a persistent launch shape (grid 1x the
    resident workgroups, four centres per edge-backward group)."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    from superlinear_amd.pair import PairKernel
    print("[case] pair persistent", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frame = W.periodic_frames(1, 100, 100, 2, 0.10, seed=61)[0]
    m = Sus2Model(path)
    m.tune("fp64", 4, 1)
    ref = oracle_lib.OracleModel(path).efs([frame])
    ij, d = oracle_lib.neighbors(frame["positions"], frame["cell"], True, m.cutoff)
    n = len(frame["types"])
    k = PairKernel(m, "fp64")
    k.begin(n, n, frame["types"], ij[:, 0], ij[:, 1], d)  # one domain owning every atom: no ghosts
    k.outer()
    r = k.finish()
    assert _maxrel(r["force"], ref["forces"]) <= 1e-10
    assert abs(r["energy"] - ref["energy"][0]) <= 1e-10 * abs(ref["energy"][0])
    assert _maxrel(r["virial"], ref["virial"][0]) <= 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_message_modes_agree(lib, oracle_lib, prec):
    """The three ways of forming the outer pass's gate-adjoint messages (one
    combined edge pass, a forces-only pass plus a messages-only pass, or
    recomputation in the gather) give the oracle's EFS; the split form equals
    the combined one bit for bit."""
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print(f"[case] message modes {prec}", flush=True)
    path = W.model_path("effect_l4k6r2o3")
    frames = _mixed_frames(3, 71)
    ref = oracle_lib.OracleModel(path).efs(frames, nthreads=16)
    m = Sus2Model(path)
    assert m.messages(prec) == "auto"
    batch = DeviceBatch(frames)
    out = {}
    for mode in ("store", "split", "recompute", "auto"):
        assert m.messages(prec, mode) == mode
        out[mode] = {k: v.cpu().numpy() for k, v in m.efs_device(batch, prec).items()}
        torch.cuda.synchronize()
    for k in ("energy", "forces", "virial", "site"):
        assert np.array_equal(out["store"][k], out["split"][k]), k
        assert np.array_equal(out["auto"][k], out["split"][k]), k  # automatic = split at L4
        assert _maxrel(out["recompute"][k], out["store"][k]) <= (1e-13 if prec == "fp64" else 1e-6), k
    tol = 1e-10 if prec == "fp64" else 1e-5
    for mode in ("store", "recompute"):
        assert _maxrel(out[mode]["forces"], ref["forces"]) <= tol
        assert _maxrel(out[mode]["energy"], ref["energy"]) <= tol
