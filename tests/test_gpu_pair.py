"""GPU checks of the pair-style compute() contract (begin -> forward comm ->
outer -> reverse comm -> finish) under an emulated domain decomposition: the
atoms of a periodic frame are split over "ranks" that each own a slab and see
everything else within the cutoff as ghosts (periodic images included).
Forward comm copies the owners' gate signals into the ghost rows, reverse comm
adds the ghosts' adjoints back to their owners, and ghost forces are summed
onto their owners, as LAMMPS does.  The assembled energy, forces and virial
must equal the oracle's periodic EFS of the whole frame."""
import numpy as np
import pytest

from conftest import close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu_available):
    from superlinear_amd import _lib
    _lib.lib()
    return _lib


def _decompose(frame, nranks, rc):
    """Per rank: owned atom ids, ghost (owner id, position) list and the owned
    atoms' full edge lists over owned + ghosts."""
    x = frame["positions"]
    cell = frame["cell"]
    n = len(x)
    inv = np.linalg.inv(cell)
    frac = (x @ inv)[:, 0]
    owner = np.minimum((np.mod(frac, 1.0) * nranks).astype(int), nranks - 1)
    # every image of every atom within reach
    reach = rc * np.linalg.norm(inv, axis=0)
    S = np.floor(reach).astype(int) + 2
    imgs = []
    for a in range(-S[0], S[0] + 1):
        for b in range(-S[1], S[1] + 1):
            for c in range(-S[2], S[2] + 1):
                k = np.array([a, b, c])
                for j in range(n):
                    imgs.append((j, x[j] + k @ cell, (a, b, c) == (0, 0, 0)))
    ranks = []
    for r in range(nranks):
        own = np.flatnonzero(owner == r)
        local = {int(j): q for q, j in enumerate(own)}
        ghosts = []  # (owner atom id, position)
        ei, ej, d = [], [], []
        gidx = {}
        for q, i in enumerate(own):
            for (j, pj, home) in imgs:
                dv = pj - x[i]
                r2 = dv @ dv
                if r2 >= rc * rc or (home and j == i):
                    continue
                if home and owner[j] == r:
                    jl = local[j]
                else:
                    key = (j, tuple(np.round(pj, 9)))
                    if key not in gidx:
                        gidx[key] = len(own) + len(ghosts)
                        ghosts.append((j, pj))
                    jl = gidx[key]
                ei.append(q)
                ej.append(jl)
                d.append(dv)
        ranks.append(dict(own=own, ghosts=ghosts, ei=np.array(ei), ej=np.array(ej), d=np.array(d).reshape(-1, 3)))
    return ranks, owner


@pytest.mark.parametrize("mname,species,nranks,prec", [("effect_l3k4r2o2", 2, 2, "fp64"),
                                                       ("effect_l3k4r2o2", 2, 1, "fp64"),
                                                       ("static_l3k4o2", 1, 2, "fp64"),
                                                       ("effect_l3k4r2o2", 2, 2, "fp32")])
def test_pair_contract_domain_decomposition(lib, oracle_lib, mname, species, nranks, prec):
    from superlinear_amd import workloads as W
    from superlinear_amd.model import Sus2Model
    from superlinear_amd.pair import PairKernel
    print(f"[case] pair {mname} ranks={nranks} {prec}", flush=True)
    path = W.model_path(mname)
    frame = W.diamond(2, species=species, seed=91)
    ref = oracle_lib.OracleModel(path).efs([frame])
    model = Sus2Model(path)
    rcut = model.cutoff
    ranks, owner = _decompose(frame, nranks, rcut)
    types = frame["types"]
    kernels = []
    for rk in ranks:
        k = PairKernel(model, prec)
        nown, ng = len(rk["own"]), len(rk["ghosts"])
        t = np.concatenate([types[rk["own"]], np.array([types[j] for j, _ in rk["ghosts"]], np.int32)])
        k.begin(nown, nown + ng, t, rk["ei"], rk["ej"], rk["d"])
        kernels.append(k)
    w = kernels[0].communication_width()
    assert w == (model.dims()["F"] if model.dims()["gated"] else 0)
    loc = [{int(j): q for q, j in enumerate(rk["own"])} for rk in ranks]
    if w:  # forward comm: owners pack, ghost holders unpack
        for r, rk in enumerate(ranks):
            if not rk["ghosts"]:
                continue
            rows = np.zeros((len(rk["ghosts"]), w))
            for o in range(nranks):  # one pack per owning rank
                sel = [g for g, (j, _) in enumerate(rk["ghosts"]) if owner[j] == o]
                if sel:
                    rows[sel] = kernels[o].pack_forward([loc[o][rk["ghosts"][g][0]] for g in sel])
            kernels[r].unpack_forward(np.arange(len(rk["own"]), len(rk["own"]) + len(rk["ghosts"])), rows)
    for k in kernels:
        k.outer()
    if w:  # reverse comm: ghost adjoints back to the owners
        for r, rk in enumerate(ranks):
            if not rk["ghosts"]:
                continue
            gl = np.arange(len(rk["own"]), len(rk["own"]) + len(rk["ghosts"]))
            rows = kernels[r].pack_reverse(gl)
            for o in range(nranks):
                sel = [g for g, (j, _) in enumerate(rk["ghosts"]) if owner[j] == o]
                if sel:
                    kernels[o].add_reverse([loc[o][rk["ghosts"][g][0]] for g in sel], rows[sel])
    E = 0.0
    F = np.zeros_like(frame["positions"])
    V = np.zeros(6)
    for r, (k, rk) in enumerate(zip(kernels, ranks)):
        out = k.finish()
        E += out["energy"]
        V += out["virial"]
        F[rk["own"]] += out["force"][:len(rk["own"])]
        for g, (j, _) in enumerate(rk["ghosts"]):
            F[j] += out["force"][len(rk["own"]) + g]
    tol = 1e-10 if prec == "fp64" else 1e-5
    close(E, ref["energy"][0], tol)
    close(F, ref["forces"], tol)
    close(V, ref["virial"][0], tol if prec == "fp64" else 1e-4)


def test_pair_contract_rejects_bad_input(lib):
    from superlinear_amd import workloads as W
    from superlinear_amd._lib import Sus2Error
    from superlinear_amd.model import Sus2Model
    from superlinear_amd.pair import PairKernel
    k = PairKernel(Sus2Model(W.model_path("effect_l3k4r2o2")), "fp64")
    with pytest.raises(Sus2Error):
        k.outer()  # before begin
    with pytest.raises(Sus2Error):
        k.begin(2, 3, [0, 1, 0], [0], [5], [[1.0, 0, 0]])  # neighbour index out of range
    with pytest.raises(Sus2Error):
        k.begin(2, 3, [0, 1, 0], [2], [0], [[1.0, 0, 0]])  # centre is not owned
    with pytest.raises(Sus2Error):
        k.begin(2, 3, [0, 1, 0], [0], [1], [[0.0, 0, 0]])  # zero-length edge
    with pytest.raises(Sus2Error):
        k.begin(2, 3, [0, 9, 0], [0], [1], [[1.0, 0, 0]])  # species out of range
