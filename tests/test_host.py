"""Host-side checks that need no GPU: the C ABI library loads and exports
every declared symbol, the model parser / coefficient layout agree with the
oracle's independent parser, save/load round trips, the topology generator
reproduces the reference fixture graphs, CFG I/O."""
import io
import re

import numpy as np
import pytest

from conftest import FIXTURES, ROOT, real_sh_scipy


def _declared_symbols():
    text = (ROOT / "include" / "sus2amd.h").read_text()
    return sorted(set(re.findall(r"\b(sus2_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_declared_symbols():
    from superlinear_amd import _lib
    L = _lib.lib()
    declared = _declared_symbols()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.SIGNATURES)
    assert L.sus2_abi_version() == 1


MODELS = [FIXTURES / "fixtures_v2" / "static.mtp", FIXTURES / "fixtures_v2" / "effect.mtp",
          FIXTURES / "fixtures_hi_l" / "effect_l3k4r3o3.mtp"]


@pytest.mark.parametrize("path", MODELS + ["static_l3k4o2", "effect_l3k4r2o2", "effect_l4k6r2o3", "effect_l5k6r2o3"],
                         ids=str)
def test_parser_layout_matches_oracle(oracle_lib, path):
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    if isinstance(path, str):
        path = W.model_path(path)
    m = Sus2Model(path)
    o = oracle_lib.OracleModel(path)
    assert np.array_equal(m.coeffs(), o.coeffs())
    off_m, off_o = m.offsets(), o.offsets()
    for k in off_o:
        assert off_m[k] == off_o[k], k
    dm, do = m.dims(), o.dims()
    for k in do:
        assert dm[k] == do[k], k


def test_save_load_round_trip(tmp_path, oracle_lib):
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    m = Sus2Model(W.model_path("effect_l3k4r2o2"))
    theta = m.coeffs()
    rng = np.random.default_rng(0)
    t2 = theta + 1e-3 * rng.normal(size=theta.size)
    m.set_coeffs(t2)
    p = tmp_path / "x.sus2"
    m.save(p)
    m2 = Sus2Model(p)
    assert np.array_equal(m2.coeffs(), t2)
    # the saved file is a valid model for the independent oracle parser too
    assert np.array_equal(oracle_lib.OracleModel(p).coeffs(), t2)


def test_invalid_models_rejected(tmp_path):
    from superlinear_amd import _lib
    from superlinear_amd.model import Sus2Model
    with pytest.raises(_lib.Sus2Error):
        Sus2Model(tmp_path / "missing.sus2")
    bad = tmp_path / "bad.sus2"
    bad.write_text((FIXTURES / "fixtures_v2" / "static.mtp").read_text().replace("RBChebyshev_sss", "RBShapeev"))
    with pytest.raises(_lib.Sus2Error) as e:
        Sus2Model(bad)
    assert e.value.code == 3
    cyc = tmp_path / "cyc.sus2"
    cyc.write_text((FIXTURES / "fixtures_v2" / "static.mtp").read_text().replace("{0, 0, 8,", "{9, 0, 8,"))
    with pytest.raises(_lib.Sus2Error):
        Sus2Model(cyc)


def _fields(text):
    out = {}
    for line in text.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


def _nums(s):
    return np.array([float(x) for x in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", s)])


GRAPH_KEYS = ["alpha_moments_count", "alpha_index_basic", "sh_products", "alpha_moment_mapping", "sh_scalar_info"]
GATE_KEYS = GRAPH_KEYS + ["two_layer_gate_k_rank_basic_begin", "outer_k_rank1_basic_begin", "outer_k_rank1_product_begin",
                          "two_layer_gate_k_rank_product_end", "two_layer_gate_scalar_moment_indices",
                          "two_layer_gate_scalar_body_orders"]


@pytest.mark.parametrize("fixture,kw,keys", [
    ("fixtures_v2/static.mtp", dict(lmax=1, kmax=2, species=2, body_order=3, outer_rank=0, rb_size=3), GRAPH_KEYS),
    ("fixtures_v2/effect.mtp", dict(lmax=1, kmax=2, species=2, body_order=3, outer_rank=2, gate=True, gate_rank=2,
                                    rb_size=3), GATE_KEYS),
    ("fixtures_hi_l/effect_l3k4r3o3.mtp", dict(lmax=3, kmax=4, species=2, body_order=3, outer_rank=3, gate=True,
                                                gate_rank=3, rb_size=3), GATE_KEYS),
])
def test_topology_reproduces_reference_graphs(fixture, kw, keys):
    """The generator re-derives the product graphs (nodes, real CG
    coefficients, scalar and segment layout) serialised in the reference's
    fixtures (dev_src/sh_model_init.cpp conventions)."""
    from superlinear_amd.topology import build_model
    a = _fields((FIXTURES / fixture).read_text())
    b = _fields(build_model(**kw))
    for k in keys:
        x, y = _nums(a[k]), _nums(b[k])
        assert x.shape == y.shape, k
        assert np.abs(x - y).max(initial=0) <= 1e-12, k


REF_TOPOLOGY = ROOT / "oracle" / "_ref" / "ref-topology"

# (lmax, kmax, body order, outer pruning, gate rank, outer rank, gate body order, gate pruning); gate rank 0 = Static
TOPOLOGY_CASES = [
    (1, 2, 4, "q-total", 0, 0, 0, ""), (3, 4, 4, "q-total", 0, 0, 0, ""), (3, 4, 4, "same-k", 0, 0, 0, ""),
    (2, 3, 5, "q-total", 0, 0, 0, ""), (2, 3, 5, "same-k", 0, 0, 0, ""), (2, 2, 6, "q-total", 0, 0, 0, ""),
    (2, 2, 6, "same-k", 0, 0, 0, ""), (6, 1, 4, "q-total", 0, 0, 0, ""), (3, 4, 4, "q-total", 0, 2, 0, ""),
    (3, 4, 4, "same-k", 2, 2, 4, "q-total"), (3, 4, 4, "q-total", 2, 2, 4, "q-total"),
    (2, 3, 5, "same-k", 2, 3, 4, "q-total"), (2, 3, 5, "same-k", 3, 2, 4, "same-k"),
    (3, 3, 5, "q-total", 2, 2, 5, "same-k"), (2, 2, 6, "same-k", 2, 2, 6, "q-total"),
    (4, 6, 4, "same-k", 2, 3, 4, "q-total"), (5, 6, 4, "same-k", 2, 3, 4, "q-total"),
]


@pytest.mark.skipif(not REF_TOPOLOGY.exists(), reason="reference generator not built (oracle/Makefile ref)")
@pytest.mark.parametrize("case", TOPOLOGY_CASES, ids=lambda c: "L%dK%d_b%d_%s_r%do%d_g%d%s" % c)
def test_topology_matches_reference_generator(case, tmp_path):
    """Graphs at body orders 4..6, both factor prunings, Static and Effect
    (ranked gate and outer graphs, same-k rank graphs bound per channel)
    equal the reference generator's (dev_src/sh_model_init.cpp, built from
    its sources into oracle/_ref/ref-topology by oracle/Makefile): nodes,
    products with their coefficients, scalar nodes and body metadata."""
    import subprocess
    from superlinear_amd.topology import build_model
    L, K, body, pr, R, O, gbody, gpr = case
    opts = {"l-max": L, "k-max": K, "body-order": body, "species-count": 2, "radial-basis-size": 3,
            "sh-factor-pruning": pr}
    kw = dict(lmax=L, kmax=K, species=2, body_order=body, outer_rank=O, outer_pruning=pr, rb_size=3)
    keys = GRAPH_KEYS[:-1]
    if O:
        opts["outer-k-rank"] = O
        keys = GRAPH_KEYS + ["outer_k_rank1_basic_begin", "outer_k_rank1_product_begin", "outer_channel_mode"]
    if R:
        opts.update({"two-layer-gate": None, "two-layer-gate-mode": "mu-scalar-full",
                     "two-layer-gate-site-mode": "neighbor", "two-layer-gate-radial-mode": "shared-radial",
                     "two-layer-gate-k-rank": R, "two-layer-gate-body-order": gbody,
                     "two-layer-gate-factor-pruning": gpr})
        kw.update(gate=True, gate_rank=R, gate_body_order=gbody, gate_pruning=gpr)
        keys = GATE_KEYS + ["outer_channel_mode", "two_layer_gate_factor_pruning"]
    out = tmp_path / "ref.sus2"
    args = [str(REF_TOPOLOGY), str(out)] + [k if v is None else f"{k}={v}" for k, v in opts.items()]
    subprocess.run(args, check=True, capture_output=True, timeout=120)
    a = _fields(out.read_text())
    b = _fields(build_model(**kw))
    for k in keys:
        if k in ("outer_channel_mode", "two_layer_gate_factor_pruning"):
            # the reference writes the channel mode only where it is not the default
            assert a.get(k, "cross-channel") == b[k], k
            continue
        x, y = _nums(a[k]), _nums(b[k])
        assert x.shape == y.shape, k
        assert np.abs(x - y).max(initial=0) <= 1e-12, k


def test_topology_rejects_gate_body_order_above_outer():
    """The reference generator refuses a gate body order above the outer one
    (dev_src/sh_model_init.cpp:1374-1380)."""
    from superlinear_amd.topology import build_model
    with pytest.raises(ValueError):
        build_model(lmax=2, kmax=3, species=2, body_order=4, outer_rank=2, gate=True, gate_rank=2, gate_body_order=5)


def test_real_cg_orthogonality():
    """Real coupling coefficients form an orthogonal transform of the
    (l1 x l2) product space (sum over L, M of C C = delta), for l1, l2 <= 6."""
    from superlinear_amd.topology import real_cg
    for l1 in range(1, 7):
        for l2 in range(1, l1 + 1):
            n = (2 * l1 + 1) * (2 * l2 + 1)
            U = np.zeros((n, n))
            row = 0
            for L in range(abs(l1 - l2), l1 + l2 + 1):
                for M in range(-L, L + 1):
                    U[row] = [real_cg(l1, a, l2, b, L, M) for a in range(-l1, l1 + 1) for b in range(-l2, l2 + 1)]
                    row += 1
            assert np.allclose(U @ U.T, np.eye(n), atol=1e-12), (l1, l2)


def test_real_cg_couples_harmonics_of_one_direction():
    """sum_{r1 r2} C(l1 r1, l2 r2 | L R) Y_{l1 r1}(u) Y_{l2 r2}(u) equals
    i^(l1+l2-L) g Y_{L R}(u) with the Gaunt factor
    g = sqrt((2l1+1)(2l2+1) / (4 pi (2L+1))) <l1 0 l2 0 | L 0> for even
    l1 + l2 - L (the phase of the real coupling, dev_src/mtpr_sh.cpp:265-326)
    and vanishes for odd l1 + l2 - L.
    Y from scipy in the kernels' real basis: this fixes every coefficient's
    sign and the real-basis phase, which orthogonality alone does not."""
    from math import pi, sqrt
    from superlinear_amd.topology import clebsch_gordan, real_cg
    rng = np.random.default_rng(8)
    us = [rng.normal(size=3) for _ in range(3)]
    Ys = [real_sh_scipy(12, u) for u in us]
    for l1 in range(1, 7):
        for l2 in range(1, l1 + 1):
            for L in range(l1 - l2, l1 + l2 + 1):
                C = np.array([[real_cg(l1, a, l2, b, L, R) for a in range(-l1, l1 + 1) for b in range(-l2, l2 + 1)]
                              for R in range(-L, L + 1)])
                g = sqrt((2 * l1 + 1) * (2 * l2 + 1) / (4 * pi * (2 * L + 1))) * clebsch_gordan(l1, 0, l2, 0, L, 0)
                for Y in Ys:
                    y1, y2 = Y[l1 * l1:(l1 + 1) ** 2], Y[l2 * l2:(l2 + 1) ** 2]
                    got = C @ np.outer(y1, y2).ravel()
                    k = l1 + l2 - L
                    want = (-1) ** (k // 2) * g * Y[L * L:(L + 1) ** 2] if k % 2 == 0 else np.zeros(2 * L + 1)
                    assert np.allclose(got, want, atol=1e-12), (l1, l2, L)


def test_cfg_round_trip(tmp_path):
    from superlinear_amd.cfg import read_cfg, write_cfg
    frames = list(read_cfg(FIXTURES / "fixtures_v2" / "geometries.cfg"))
    assert len(frames) == 7
    f = dict(frames[0], forces=np.arange(9.0).reshape(3, 3) * 0.1, energy=-1.5, virial=np.arange(6.0))
    buf = io.StringIO()
    write_cfg(buf, f, lossless=True)
    path = tmp_path / "rt.cfg"
    path.write_text(buf.getvalue())
    g = next(read_cfg(path))
    assert np.array_equal(g["positions"], f["positions"])
    assert np.array_equal(g["forces"], f["forces"])
    assert g["energy"] == f["energy"]
    assert np.array_equal(g["virial"], f["virial"])
    mol = dict(types=np.array([0, 1]), positions=np.array([[0, 0, 0], [1.0, 0, 0]]), periodic=False)
    buf = io.StringIO()
    write_cfg(buf, mol)
    assert "Supercell" not in buf.getvalue()


def test_cfg_partial_supercell_round_trip(tmp_path):
    """One or two Supercell rows (or a zero row) leave the other directions
    non-periodic; the writer emits rows up to the last non-zero one."""
    from superlinear_amd.cfg import read_cfg, write_cfg
    from superlinear_amd import workloads as W
    frames = W.partial_frames(1, 5, 7, 2, seed=3, axes=2) + W.partial_frames(1, 5, 7, 2, seed=4, axes=1)
    p = tmp_path / "p.cfg"
    with open(p, "w") as fh:
        for f in frames:
            write_cfg(fh, f, lossless=True)
    text = p.read_text()
    assert text.count("Supercell") == 2
    back = list(read_cfg(p))
    for f, g in zip(frames, back):
        assert g["periodic"]
        np.testing.assert_array_equal(g["cell"], f["cell"])
        np.testing.assert_array_equal(g["positions"], f["positions"])


def test_calculator_validates_before_evaluation():
    """Species mapping and geometry checks of the Python entry point
    (interfaces/python/bindings/core.cpp:26-74) raise before any GPU call."""
    from superlinear_amd import workloads as W
    from superlinear_amd.calculator import SuperlinearCalculator, SuperlinearModel
    path = W.model_path("effect_l3k4r2o2")
    with pytest.raises(ValueError):
        SuperlinearModel(path, [6])  # the model has two species
    with pytest.raises(ValueError):
        SuperlinearModel(path, [6, 6])
    with pytest.raises(ValueError):
        SuperlinearModel(path, [])
    m = SuperlinearModel(path, [6, 14])
    assert m.metadata["species_count"] == 2 and m.metadata["kind"] == "coordination-effect"
    cell = np.eye(3) * 10
    with pytest.raises(ValueError):
        m.calculate([6, 7], np.zeros((2, 3)), cell, [True] * 3)  # element outside the mapping
    with pytest.raises(ValueError):
        m.calculate([6], np.zeros((2, 3)), cell, [True] * 3)  # shape mismatch
    with pytest.raises(ValueError):
        m.calculate([6, 6], np.zeros((2, 3)), np.diag([10.0, 10.0, 0.2]), [True] * 3)  # thin periodic cell
    with pytest.raises(ValueError):
        SuperlinearCalculator(path, "CSi")
    with pytest.raises(ValueError):
        SuperlinearCalculator(path, ["C", "Xx"])


@pytest.fixture(scope="module")
def optim_driver(tmp_path_factory):
    import subprocess
    exe = tmp_path_factory.mktemp("optim") / "optim_driver"
    subprocess.run(["g++", "-O2", "-std=c++17", str(ROOT / "tests" / "native" / "optim_driver.cpp"), "-o", str(exe)],
                   check=True)
    return exe


def _run_optim(exe, *args):
    import subprocess
    out = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, check=True).stdout.split()
    head = dict(zip(out[0:8:2], out[1:8:2]))
    return head, np.array([float(v) for v in out[8:]])


def test_bfgs_quadratic_minimiser_and_secant(optim_driver, tmp_path):
    """Host BFGS on a convex quadratic with condition number 1e3: converges to
    A^-1 b, and every inverse-Hessian update satisfies the secant equation
    H y = s (the defining property of the BFGS update)."""
    rng = np.random.default_rng(17)
    n = 12
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = q @ np.diag(np.logspace(0, 3, n)) @ q.T
    b = rng.normal(size=n)
    f = tmp_path / "quad.txt"
    f.write_text(f"{n}\n" + " ".join(repr(float(v)) for v in A.ravel()) + "\n" + " ".join(repr(float(v)) for v in b))
    head, x = _run_optim(optim_driver, "quad", f)
    xs = np.linalg.solve(A, b)
    assert head["converged"] == "1"
    assert int(head["iters"]) <= 3 * n
    assert np.abs(x - xs).max() <= 1e-8 * (1 + np.abs(xs).max())
    assert float(head["secant"]) <= 1e-9 * (1 + np.abs(xs).max())


def test_bfgs_rosenbrock(optim_driver):
    head, x = _run_optim(optim_driver, "rosen")
    assert head["converged"] == "1"
    assert np.allclose(x, [1.0, 1.0], atol=1e-8)
    assert int(head["iters"]) < 100


def test_lbfgs_quadratic_and_rosenbrock(optim_driver, tmp_path):
    """L-BFGS (reference update rules) on convex quadratics: with condition
    number 10 it converges (max |g| <= 1e-9) to A^-1 b for a short and a full
    history; with condition number 1e3, where the reference's damping against
    B0 = I / gamma also fires on convex steps, every accepted step still lowers
    f and the iterate gets within 1e-6 of A^-1 b.  It minimises the
    Rosenbrock function."""
    n = 12
    for cond, hists, tol, conv in ((1, (3, 20), 1e-8, True), (3, (20,), 1e-6, False)):
        rng = np.random.default_rng(19)
        q, _ = np.linalg.qr(rng.normal(size=(n, n)))
        A = q @ np.diag(np.logspace(0, cond, n)) @ q.T
        b = rng.normal(size=n)
        f = tmp_path / f"quad{cond}.txt"
        f.write_text(f"{n}\n" + " ".join(repr(float(v)) for v in A.ravel()) + "\n" +
                     " ".join(repr(float(v)) for v in b))
        xs = np.linalg.solve(A, b)
        for hist in hists:
            head, x = _run_optim(optim_driver, "lbfgs-quad", f, hist)  # exit status 3 if f ever rose
            if conv:
                assert head["converged"] == "1" and head["damped"] == "0"
            assert np.abs(x - xs).max() <= tol * (1 + np.abs(xs).max())
    head, x = _run_optim(optim_driver, "lbfgs-rosen")
    assert head["converged"] == "1"
    assert np.allclose(x, [1.0, 1.0], atol=1e-7)


def test_lbfgs_gamma_and_powell_damping(optim_driver):
    """H0 = gamma I with gamma = s.y / y.y: on f = c x^2 / 2 (c = 1/2) the
    first pair gives gamma = 1/c and the next unit step lands on the minimum.
    From the concave middle of a double well the first pair has s.y < 0: it is
    Powell-damped (not dropped) and the iteration still reaches a minimum
    (|x| = 1, y = 0)."""
    head, x = _run_optim(optim_driver, "lbfgs-1d")
    assert float(head["gamma"]) == pytest.approx(2.0, rel=1e-14)
    assert head["converged"] == "1" and int(head["iters"]) == 2 and x[0] == 0.0
    head, x = _run_optim(optim_driver, "lbfgs-dw")
    assert int(head["damped"]) >= 1
    assert head["converged"] == "1"
    assert abs(abs(x[0]) - 1.0) <= 1e-9 and abs(x[1]) <= 1e-9


def test_shape_build_prebuilt_cached_and_limits(monkeypatch):
    """sus2_shape_build: prebuilt shapes need no shape library (""), the
    shapes compiled by build() are found in the cache (SUS2_JIT=0: no compiler
    run here), and shapes beyond the kernels' limits or absent from the cache
    with compilation disabled fail with SUS2_E_UNSUPPORTED."""
    import ctypes
    import os
    from superlinear_amd import _lib
    from superlinear_amd.build import EXTRA_SHAPES
    L = _lib.lib()
    monkeypatch.setenv("SUS2_JIT", "0")
    buf = ctypes.create_string_buffer(4096)
    assert L.sus2_shape_build(3, 4, 10, 2, buf, len(buf)) == 0 and buf.value == b""
    for shape in EXTRA_SHAPES:
        assert L.sus2_shape_build(*shape, buf, len(buf)) == 0, L.sus2_last_error()
        p = buf.value.decode()
        assert os.path.exists(p) and "sus2shape_l{}k{}b{}m{}_".format(*shape) in p
    assert L.sus2_shape_build(7, 2, 8, 2, buf, len(buf)) == 3  # L > 6
    assert L.sus2_shape_build(6, 10, 8, 2, buf, len(buf)) == 3  # K (L + 1) > 64
    assert L.sus2_shape_build(2, 3, 17, 2, buf, len(buf)) == 3  # B > 16
    assert L.sus2_shape_build(2, 3, 7, 1, buf, len(buf)) == 3  # not cached, compiler disabled
    assert b"SUS2_JIT=0" in L.sus2_last_error()
