"""GPU parity of the HIP EFS path against the FP64 oracle and the reference's
golden fixtures (the HIP path is called through the C ABI)."""
import numpy as np
import pytest

from conftest import close, fixture_cases, voigt_to_matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_available):
    from superlinear_amd import _lib
    _lib.lib()
    return _lib


@pytest.mark.parametrize("case", fixture_cases(), ids=lambda c: c[0])
def test_fixture_fp64(lib, case):
    """FP64 HIP EFS matches the reference's FP64 fixtures to 1e-9(1+|x|)
    (tolerance of embedded_two_cutoff_check.py:146) and the oracle to 1e-10."""
    from superlinear_amd.model import Sus2Model
    name, path, frame, rec = case
    print(f"[case] fixture {name}", flush=True)
    r = Sus2Model(path).calc_efs([frame], "fp64")
    close(r["energy"][0], rec["energy"], 1e-9)
    close(r["forces"], rec["forces"], 1e-9)
    close(voigt_to_matrix(r["virial"][0]), rec["plus_stress"], 1e-9)
    close(r["site"].sum(), r["energy"][0], 1e-12)
    close(r["forces"].sum(0), np.zeros(3), 1e-10)


def _oracle_efs(oracle_lib, path, frames):
    return oracle_lib.OracleModel(path).efs(frames, nthreads=8)


CONFIG_CASES = [
    ("static_l3k4o2", "diamond2", dict(ncell=2, species=1, seed=7)),
    ("effect_l3k4r2o2", "diamond2", dict(ncell=2, species=2, seed=8)),
    ("effect_l3k4r2o2", "diamond3", dict(ncell=3, species=2, seed=9)),
    ("effect_l4k6r2o3", "diamond2", dict(ncell=2, species=3, seed=10)),
    ("effect_l5k6r2o3", "diamond2", dict(ncell=2, species=2, seed=11)),
]


@pytest.mark.parametrize("mname,label,kw", CONFIG_CASES, ids=[f"{c[0]}-{c[1]}" for c in CONFIG_CASES])
def test_config_models_fp64_fp32(lib, oracle_lib, mname, label, kw):
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print(f"[case] {mname} {label}", flush=True)
    path = W.model_path(mname)
    frames = [W.diamond(**kw)]
    ref = _oracle_efs(oracle_lib, path, frames)
    m = Sus2Model(path)
    r64 = m.calc_efs(frames, "fp64")
    scale = 1 + np.abs(ref["forces"]).max()
    assert abs(r64["energy"][0] - ref["energy"][0]) <= 1e-10 * (1 + abs(ref["energy"][0]))
    assert np.abs(r64["forces"] - ref["forces"]).max() <= 1e-10 * scale
    assert np.abs(r64["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max())
    r32 = m.calc_efs(frames, "fp32")
    n = len(frames[0]["types"])
    # FP32 bulk arithmetic: force max-abs error <= 1e-5 eV/A, E/atom <= 1e-6 eV
    assert np.abs(r32["forces"] - ref["forces"]).max() <= 1e-5
    assert abs(r32["energy"][0] - ref["energy"][0]) / n <= 1e-6
    assert np.abs(r32["virial"] - ref["virial"]).max() <= 1e-4


def test_batch_mixed_frames(lib, oracle_lib):
    """Uneven frames (periodic triclinic and molecules) in one batch."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] mixed batch", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = W.periodic_frames(3, 8, 40, 2, 0.15, seed=5) + W.molecule_frames(2, 16, 30, 2, seed=6)
    ref = _oracle_efs(oracle_lib, path, frames)
    r = Sus2Model(path).calc_efs(frames, "fp64")
    assert np.abs(r["energy"] - ref["energy"]).max() <= 1e-10 * (1 + np.abs(ref["energy"]).max())
    assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-10 * (1 + np.abs(ref["forces"]).max())
    assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max())


def test_batch_uneven_l4k6(lib, oracle_lib):
    """The batched-EFS workload's form at a reduced count: three species,
    L4/K6, log-uniform frame sizes, periodic frames and molecules mixed in one
    call, FP64 and FP32 against the oracle."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] uneven batch l4k6", flush=True)
    path = W.model_path("effect_l4k6r2o3")
    frames = (W.periodic_frames(4, 16, 160, 3, 0.10, seed=41, log_uniform=True)
              + W.molecule_frames(4, 16, 160, 3, seed=42))
    ref = _oracle_efs(oracle_lib, path, frames)
    m = Sus2Model(path)
    r = m.calc_efs(frames, "fp64")
    assert np.abs(r["energy"] - ref["energy"]).max() <= 1e-10 * (1 + np.abs(ref["energy"]).max())
    assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-10 * (1 + np.abs(ref["forces"]).max())
    assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max())
    r32 = m.calc_efs(frames, "fp32")
    counts = np.array([len(f["types"]) for f in frames])
    assert np.abs(r32["forces"] - ref["forces"]).max() <= 1e-5
    assert np.abs((r32["energy"] - ref["energy"]) / counts).max() <= 1e-6


def test_batch_of_many_frames(lib, oracle_lib):
    """A batch of 48 periodic frames (more than 16384 atoms, several
    thousand neighbour-search workgroups): FP64 and FP32 against the oracle."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] batch of many frames", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = W.periodic_frames(48, 320, 440, 2, 0.05, seed=51)
    assert sum(len(f["types"]) for f in frames) > 16384
    ref = _oracle_efs(oracle_lib, path, frames)
    m = Sus2Model(path)
    r = m.calc_efs(frames, "fp64")
    assert np.abs(r["energy"] - ref["energy"]).max() <= 1e-10 * (1 + np.abs(ref["energy"]).max())
    assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-10 * (1 + np.abs(ref["forces"]).max())
    assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max())
    r32 = m.calc_efs(frames, "fp32")
    assert np.abs(r32["forces"] - ref["forces"]).max() <= 1e-5


def test_large_frame_grid_path(lib, oracle_lib):
    """A frame above the per-frame single-workgroup cell-list limit (8192
    atoms) takes the device-wide binning path; sparse (density 0.03 A^-3) so
    that the oracle stays quick."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] large frame grid path", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = W.periodic_frames(1, 8300, 8300, 2, 0.10, seed=13, density=(0.03, 0.03))
    ref = _oracle_efs(oracle_lib, path, frames)
    r = Sus2Model(path).calc_efs(frames, "fp64")
    assert abs(r["energy"][0] - ref["energy"][0]) <= 1e-10 * (1 + abs(ref["energy"][0]))
    assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-10 * (1 + np.abs(ref["forces"]).max())
    assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max())


def test_large_open_frames_grid_path(lib, oracle_lib):
    """A 9000-atom molecule and an 8600-atom slab (periodic along two lattice
    rows, the third zero): both above the per-frame cell-list limit, so the
    device-wide binning with open-direction extents builds their lists.
    FP64 and FP32 against the oracle."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] large open frames grid path", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = (W.molecule_frames(1, 9000, 9000, 2, seed=61, density=0.04)
              + W.partial_frames(1, 8600, 8600, 2, seed=62, axes=2, density=0.04))
    ref = _oracle_efs(oracle_lib, path, frames)
    m = Sus2Model(path)
    r = m.calc_efs(frames, "fp64")
    assert np.abs(r["energy"] - ref["energy"]).max() <= 1e-10 * (1 + np.abs(ref["energy"]).max())
    assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-10 * (1 + np.abs(ref["forces"]).max())
    assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max())
    r32 = m.calc_efs(frames, "fp32")
    assert np.abs(r32["forces"] - ref["forces"]).max() <= 1e-5


def test_neighbor_slot_overflow_then_adapted(lib, oracle_lib):
    """A dense frame (about 180 neighbours per atom) overflows the default
    128 per-atom neighbour slots: the overflowed atoms are filled directly and
    the incoming lists built afterwards.  The next call uses the enlarged
    slots (copy and incoming placement fused).  Both against the oracle."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] neighbour slot overflow", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = W.periodic_frames(1, 64, 64, 2, 0.05, seed=17, density=(0.35, 0.35))
    ref = _oracle_efs(oracle_lib, path, frames)
    m = Sus2Model(path)
    for call in range(2):
        r = m.calc_efs(frames, "fp64")
        assert abs(r["energy"][0] - ref["energy"][0]) <= 1e-10 * (1 + abs(ref["energy"][0])), call
        assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-10 * (1 + np.abs(ref["forces"]).max()), call
        assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max()), call


def _directed_pairs(frame, rc):
    """Brute-force count of directed pairs closer than rc over periodic images."""
    x = np.asarray(frame["positions"], np.float64)
    a = np.asarray(frame["cell"], np.float64)
    recip = np.linalg.inv(a).T  # rows: reciprocal vectors (a_q . b_p = delta)
    n = [int(np.ceil(rc * np.linalg.norm(recip[q]))) + 1 for q in range(3)]
    d = x[None, :, :] - x[:, None, :]
    count = 0
    for s0 in range(-n[0], n[0] + 1):
        for s1 in range(-n[1], n[1] + 1):
            for s2 in range(-n[2], n[2] + 1):
                r2 = ((d + s0 * a[0] + s1 * a[1] + s2 * a[2]) ** 2).sum(-1)
                hit = r2 < rc * rc
                if s0 == s1 == s2 == 0:
                    np.fill_diagonal(hit, False)
                count += int(hit.sum())
    return count


def test_edge_capacity_from_earlier_calls(lib, oracle_lib):
    """Calls after the first size the neighbour lists from earlier calls and
    read the device's overflow report at the end.  A sparse frame sets the
    capacity; a dense frame then overflows the slots and the edge buffers, a
    larger frame of normal density the edge buffers only, and a small frame
    fits.  Each call's EFS and edge count against the oracle."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] edge capacity from earlier calls", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    seq = [W.periodic_frames(1, 48, 48, 2, 0.05, seed=31, density=(0.08, 0.08)),
           W.periodic_frames(1, 64, 64, 2, 0.05, seed=32, density=(0.35, 0.35)),
           W.periodic_frames(1, 400, 400, 2, 0.05, seed=33),
           W.periodic_frames(1, 40, 40, 2, 0.05, seed=34)]
    m = Sus2Model(path)
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        for call, frames in enumerate(seq):
            ref = _oracle_efs(oracle_lib, path, frames)
            r = m.calc_efs(frames, prec)
            assert m.last_edge_count() == _directed_pairs(frames[0], m.cutoff), (prec, call)
            assert abs(r["energy"][0] - ref["energy"][0]) <= tol * (1 + abs(ref["energy"][0])), (prec, call)
            assert np.abs(r["forces"] - ref["forces"]).max() <= tol * (1 + np.abs(ref["forces"]).max()), (prec, call)
            assert np.abs(r["virial"] - ref["virial"]).max() <= tol * (1 + np.abs(ref["virial"]).max()), (prec, call)


def test_cutoff_ties_keep_lists_symmetric(lib, oracle_lib):
    """Simple cubic lattice with spacing rc/3: many separations round to within
    an ulp of the cutoff.  Both directions of a pair must pass the cutoff test
    together (the incoming lists rely on a symmetric directed list): the edge
    count is even and the EFS matches the oracle (the smooth cutoff makes the
    contributions of pairs at rc vanish, so the oracle's own tie decisions do
    not matter)."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] cutoff ties", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    m = Sus2Model(path)
    a = m.cutoff / 3.0
    n = 6
    grid = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(23)
    frames = [dict(types=rng.integers(0, 2, n ** 3).astype(np.int32), positions=grid * a + 0.3,
                   cell=np.eye(3) * (n * a), periodic=True)]
    ref = _oracle_efs(oracle_lib, path, frames)
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        r = m.calc_efs(frames, prec)
        assert m.last_edge_count() % 2 == 0, prec
        assert abs(r["energy"][0] - ref["energy"][0]) <= tol * (1 + abs(ref["energy"][0])), prec
        assert np.abs(r["forces"] - ref["forces"]).max() <= tol * (1 + np.abs(ref["forces"]).max()), prec


def test_atoms_far_outside_the_cell(lib, oracle_lib):
    """Atoms moved by whole lattice vectors far outside the cell (wraps up to
    60 cells, beyond what the per-frame binning keeps in its LDS bin word):
    the same EFS as the wrapped frame and the oracle."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] atoms far outside the cell", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frame = W.diamond(ncell=2, species=2, seed=12)
    a = np.asarray(frame["cell"], np.float64)
    rng = np.random.default_rng(5)
    shifts = rng.integers(-60, 61, (len(frame["types"]), 3))
    shifts[::3] = 0
    far = dict(frame, positions=np.asarray(frame["positions"]) + shifts @ a)
    ref = _oracle_efs(oracle_lib, path, [frame])
    m = Sus2Model(path)
    for fr in (frame, far):
        r = m.calc_efs([fr], "fp64")
        assert abs(r["energy"][0] - ref["energy"][0]) <= 1e-9 * (1 + abs(ref["energy"][0]))
        assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-9 * (1 + np.abs(ref["forces"]).max())
        assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-9 * (1 + np.abs(ref["virial"]).max())


def test_tiny_cells_wide_search(lib, oracle_lib):
    """Periodic cells of one and two atoms, 2.2-2.6 A wide (one triclinic): the
    5 A cutoff reaches three cells out, a 7 x 7 x 7 bin search whose incoming
    positions come from a binary search over each target's hit offsets.
    FP64 and FP32 against the oracle."""
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    print("[case] tiny cells", flush=True)
    path = W.model_path("effect_l3k4r2o2")
    frames = [dict(types=np.array([0], np.int32), positions=np.array([[0.1, 0.2, 0.3]]),
                   cell=np.eye(3) * 2.2, periodic=True),
              dict(types=np.array([0, 1], np.int32), positions=np.array([[0.0, 0.0, 0.0], [1.2, 1.1, 1.3]]),
                   cell=np.array([[2.6, 0.0, 0.0], [0.7, 2.4, 0.0], [0.3, 0.5, 2.5]]), periodic=True)]
    ref = _oracle_efs(oracle_lib, path, frames)
    m = Sus2Model(path)
    r = m.calc_efs(frames, "fp64")
    assert np.abs(r["energy"] - ref["energy"]).max() <= 1e-10 * (1 + np.abs(ref["energy"]).max())
    assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-10 * (1 + np.abs(ref["forces"]).max())
    assert np.abs(r["virial"] - ref["virial"]).max() <= 1e-10 * (1 + np.abs(ref["virial"]).max())
    r32 = m.calc_efs(frames, "fp32")
    assert np.abs(r32["forces"] - ref["forces"]).max() <= 1e-5


def test_device_entry_point_matches_host(lib):
    import torch
    from superlinear_amd import workloads as W
    from superlinear_amd.model import DeviceBatch, Sus2Model
    print("[case] device entry", flush=True)
    m = Sus2Model(W.model_path("effect_l3k4r2o2"))
    frames = [W.diamond(2, species=2, seed=3), W.diamond(2, species=2, seed=4)]
    host = m.calc_efs(frames, "fp64")
    db = DeviceBatch(frames)
    out = m.efs_device(db, "fp64")
    torch.cuda.synchronize()
    assert np.array_equal(out["energy"].cpu().numpy(), host["energy"])
    assert np.array_equal(out["forces"].cpu().numpy(), host["forces"])
