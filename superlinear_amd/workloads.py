"""Synthetic workloads: seeded models and frames for the benchmark configs.

Five configurations of the SUS2-SH EFS and loss-gradient path:

1. ``static64``: Static L3/K4/O2 (outer coupling cross), 1 species, 64-atom
   diamond (2x2x2 conventional cells, a = 3.567 A), Gaussian displacements
   sigma = 0.10 A, cutoff 5.0 A.
2. ``effect4096``: Effect L3/K4/R2/O2 (gate coupling cross, outer coupling
   diagonal), 2 species drawn 50/50, 4096-atom diamond (8x8x8), sigma 0.10 A.
3. ``train256``: Effect L3/K4/R2/O2, 2 species, 256 periodic frames with
   N ~ U{8..216}, random triclinic cells of density 0.17-0.19 A^-3, sigma
   0.15 A; labels from a seeded teacher model of the same form.
4. ``batch4096``: Effect L4/K6/R2/O3, 3 species, 4096 frames with
   N ~ logU[16, 512], half periodic cells and half molecules.
5. ``md2m``: Effect L5/K6/R2/O3, 2 species, diamond 64^3 cells
   (2,097,152 atoms), sigma 0.10 A.

Common model defaults (the reference's example/train-effect.yaml): radial
basis size 10, warp sharing per radial channel K (scaling_map = K), body
order 4 for both graphs, per-element readout, basis scale 0.01, radial
coefficients ~ N(0, 0.1^2), gate weights ~ N(0, 0.1^2), alpha = 1/12, gate
type coefficients 1, rank weights with orthonormal rows.
"""
import os

import numpy as np

from .topology import build_model

A_DIAMOND = 3.567
CUTOFF = 5.0

MODELS = {
    "static_l3k4o2": dict(lmax=3, kmax=4, species=1, body_order=4, outer_rank=2, outer_pruning="q-total",
                          gate=False, seed=11),
    "effect_l3k4r2o2": dict(lmax=3, kmax=4, species=2, body_order=4, outer_rank=2, outer_pruning="same-k",
                            gate=True, gate_rank=2, gate_pruning="q-total", seed=22),
    "teacher_l3k4r2o2": dict(lmax=3, kmax=4, species=2, body_order=4, outer_rank=2, outer_pruning="same-k",
                             gate=True, gate_rank=2, gate_pruning="q-total", seed=33),
    "effect_l4k6r2o3": dict(lmax=4, kmax=6, species=3, body_order=4, outer_rank=3, outer_pruning="same-k",
                            gate=True, gate_rank=2, gate_pruning="q-total", seed=44),
    "effect_l5k6r2o3": dict(lmax=5, kmax=6, species=2, body_order=4, outer_rank=3, outer_pruning="same-k",
                            gate=True, gate_rank=2, gate_pruning="q-total", seed=55),
}

CONFIG_MODEL = {"static64": "static_l3k4o2", "effect4096": "effect_l3k4r2o2", "train256": "effect_l3k4r2o2",
                "batch4096": "effect_l4k6r2o3", "md2m": "effect_l5k6r2o3"}


def model_path(name, directory=None):
    """Return the path of a named synthetic model, (re)written whenever the
    cached file differs from what the generator gives now."""
    directory = directory or os.environ.get("SUS2_MODEL_DIR") or os.path.join(
        os.path.dirname(os.path.abspath(__file__)), "_models")
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, name + ".sus2")
    text = build_model(cutoff=CUTOFF, name=name, **MODELS[name])
    if os.path.exists(path):
        with open(path) as fh:
            if fh.read() == text:
                return path
    tmp = path + f".tmp{os.getpid()}"
    with open(tmp, "w") as fh:
        fh.write(text)
    os.replace(tmp, path)
    return path


def diamond(ncell, a=A_DIAMOND, sigma=0.10, species=1, seed=0):
    """Diamond lattice of ncell^3 conventional cells (8 atoms each) with
    Gaussian displacements; species drawn uniformly when species > 1."""
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    g = np.arange(ncell)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    frac = (cells[:, None, :] + base[None, :, :]).reshape(-1, 3)
    x = frac * a + sigma * rng.standard_normal(frac.shape)
    n = len(x)
    types = rng.integers(0, species, n).astype(np.int32) if species > 1 else np.zeros(n, np.int32)
    return dict(types=types, positions=x, cell=np.eye(3) * a * ncell, periodic=True)


def random_cell(n, density, rng):
    """A random, well-conditioned triclinic cell of volume n / density."""
    vol = n / density
    while True:
        m = np.eye(3) + 0.25 * rng.uniform(-1, 1, (3, 3))
        det = np.linalg.det(m)
        if det > 0.5:
            cell = m * (vol / det) ** (1.0 / 3.0)
            # reject thin cells: every face distance at least 60% of the cube edge
            inv = np.linalg.inv(cell)
            heights = 1.0 / np.linalg.norm(inv, axis=0)
            if heights.min() > 0.6 * vol ** (1.0 / 3.0):
                return cell


def random_positions(n, cell, rng, min_sep=1.0, periodic=True, max_tries=200000):
    """Uniform positions with rejection of candidates closer than min_sep to an
    accepted point (minimum image for periodic cells)."""
    pts = np.zeros((n, 3))
    inv = np.linalg.inv(cell)
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms with the requested separation")
        p = rng.uniform(0, 1, 3) @ cell
        if k:
            d = p - pts[:k]
            if periodic:
                f = d @ inv
                f -= np.round(f)
                d = f @ cell
            if np.any(np.einsum("ij,ij->i", d, d) < min_sep * min_sep):
                continue
        pts[k] = p
        k += 1
    return pts


def periodic_frames(count, nmin, nmax, species, sigma, seed, density=(0.17, 0.19), log_uniform=False):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if log_uniform:
            n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        else:
            n = int(rng.integers(nmin, nmax + 1))
        cell = random_cell(n, rng.uniform(*density), rng)
        x = random_positions(n, cell, rng) + sigma * rng.standard_normal((n, 3))
        out.append(dict(types=rng.integers(0, species, n).astype(np.int32), positions=x, cell=cell, periodic=True))
    return out


def molecule_frames(count, nmin, nmax, species, seed, density=0.1):
    """Non-periodic random clusters with minimum separation 1.0 A."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        edge = (n / density) ** (1.0 / 3.0)
        x = random_positions(n, np.eye(3) * edge, rng, periodic=False)
        out.append(dict(types=rng.integers(0, species, n).astype(np.int32), positions=x, cell=np.zeros((3, 3)),
                        periodic=False))
    return out


def partial_frames(count, nmin, nmax, species, seed, axes=2, density=0.17, sigma=0.10):
    """Frames periodic along `axes` (1 or 2) lattice rows only; the remaining
    rows are zero (the reference's marking of a non-periodic direction), with
    the atoms in a finite slab (axes=2) or wire (axes=1) along them."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        cell = random_cell(n, density, rng)
        x = random_positions(n, cell, rng) + sigma * rng.standard_normal((n, 3))
        cell = cell.copy()
        cell[axes:] = 0.0
        out.append(dict(types=rng.integers(0, species, n).astype(np.int32), positions=x, cell=cell, periodic=True))
    return out


def config_frames(config, scale=1.0):
    """Frames of a named config (`scale` < 1 shrinks frame counts for tests)."""
    if config == "static64":
        return [diamond(2, sigma=0.10, species=1, seed=101)]
    if config == "effect4096":
        return [diamond(8, sigma=0.10, species=2, seed=202)]
    if config == "train256":
        return periodic_frames(max(1, int(256 * scale)), 8, 216, 2, 0.15, seed=303)
    if config == "batch4096":
        k = max(2, int(4096 * scale))
        return (periodic_frames(k // 2, 16, 512, 3, 0.10, seed=404, log_uniform=True)
                + molecule_frames(k - k // 2, 16, 512, 3, seed=405))
    if config == "md2m":
        n = max(2, int(round(64 * scale ** (1 / 3))))
        return [diamond(n, sigma=0.10, species=2, seed=505)]
    raise KeyError(config)
