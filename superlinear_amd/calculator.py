"""Python / ASE entry point of the MI355X path.

``SuperlinearModel(model, atomic_numbers)`` mirrors the reference's native
binding (interfaces/python/bindings/core.cpp:20-147): an explicit species
mapping (model type index -> atomic number), ``calculate(numbers, positions,
cell, pbc)`` returning energy, forces, the PlusStress virial as a 3x3 matrix,
the cell volume and the stress (-virial / volume in xx yy zz yz xz xy order,
or None for a zero-volume cell).  Non-periodic directions are zero lattice
rows, as in the reference's Configuration.  Evaluation runs on the GPU through
the C ABI (FP64 by default, FP32 on request).

``SuperlinearCalculator(model, species)`` mirrors the reference's ASE
calculator (interfaces/python/pysuperlinear/calculator.py:10-73); it derives
from ase.calculators.calculator.Calculator when ASE is importable and is a
plain object with the same ``calculate``/``results`` contract otherwise.
"""
from pathlib import Path

import numpy as np

from .model import Sus2Model

# chemical symbols by atomic number (index 0 unused)
_SYMBOLS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
    "Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
ATOMIC_NUMBERS = {s: z for z, s in enumerate(_SYMBOLS) if z > 0}


def _check_finite(x, what):
    if not np.all(np.isfinite(x)):
        raise ValueError(f"nonfinite {what}")


class SuperlinearModel:
    """Persistent model handle on the GPU with an explicit species mapping."""

    def __init__(self, model, atomic_numbers, precision="fp64"):
        species = [int(z) for z in atomic_numbers]
        if not species:
            raise ValueError("explicit species mapping is required")
        if any(z < 1 or z > 118 for z in species) or len(set(species)) != len(species):
            raise ValueError("invalid or duplicate atomic number in species mapping")
        self.path = str(model)
        self.species = species
        self.types = {z: k for k, z in enumerate(species)}
        self.precision = precision
        self.model = Sus2Model(self.path)
        dims = self.model.dims()
        if dims["C"] != len(species):
            raise ValueError("species mapping length differs from model species_count")
        self._dims = dims

    def calculate(self, numbers, positions, cell, pbc):
        numbers = np.asarray(numbers)
        positions = np.asarray(positions, dtype=np.float64)
        cell = np.asarray(cell, dtype=np.float64)
        pbc = np.asarray(pbc, dtype=bool)
        if (numbers.ndim != 1 or positions.ndim != 2 or positions.shape[1:] != (3,)
                or numbers.shape[0] != positions.shape[0] or numbers.shape[0] == 0):
            raise ValueError("expected nonempty numbers[N] and positions[N,3]")
        if cell.shape != (3, 3) or pbc.shape != (3,):
            raise ValueError("expected cell[3,3] and pbc[3]")
        _check_finite(cell, "geometry or model output")
        _check_finite(positions, "geometry or model output")
        lattice = np.where(pbc[:, None], cell, 0.0)
        basis = []
        for a in range(3):  # Gram-Schmidt over the periodic rows (core.cpp:57-74)
            if not pbc[a]:
                continue
            v = cell[a].copy()
            for b in basis:
                v -= (v @ b) * b
            nrm = np.linalg.norm(v)
            if nrm < 0.5:
                raise ValueError("periodic cell is degenerate or thinner than 0.5 A")
            basis.append(v / nrm)
        try:
            types = np.array([self.types[int(z)] for z in numbers], dtype=np.int32)
        except KeyError:
            raise ValueError("element is outside the model species mapping") from None
        frame = dict(types=types, positions=positions, cell=lattice, periodic=bool(pbc.any()))
        r = self.model.calc_efs([frame], self.precision)
        energy = float(r["energy"][0])
        forces = r["forces"]
        w = r["virial"][0]  # xx yy zz xy xz yz
        virial = np.array([[w[0], w[3], w[4]], [w[3], w[1], w[5]], [w[4], w[5], w[2]]])
        _check_finite(energy, "geometry or model output")
        _check_finite(forces, "geometry or model output")
        volume = abs(float(np.linalg.det(cell)))
        stress = None
        if volume > 1e-6:
            stress = -np.array([virial[0, 0], virial[1, 1], virial[2, 2], virial[1, 2], virial[0, 2],
                                virial[0, 1]]) / volume
        return dict(energy=energy, forces=forces, virial=virial, volume=volume, stress=stress)

    @property
    def metadata(self):
        d = self._dims
        gated = bool(d["gated"])
        return dict(model_path=self.path, atomic_numbers=list(self.species), precision=self.precision,
                    backend="superlinear-amd-hip", is_sh_potential=True, species_count=d["C"],
                    cutoff=self.model.cutoff, L=d["L"], K=d["K"], coefficient_count=self.model.coeff_count,
                    basis_size=d["B"], basis_type="RBChebyshev_sss", gate_rank=d["R"], outer_rank=d["O"],
                    readout="species" if d["species_readout"] else "shared", coordination_enabled=gated,
                    kind="coordination-effect" if gated else "static-radial")


try:  # ASE is optional
    from ase.calculators.calculator import Calculator as _Base
    from ase.calculators.calculator import PropertyNotImplementedError, all_changes
    _HAVE_ASE = True
except ImportError:  # pragma: no cover - exercised where ASE is absent
    _HAVE_ASE = False
    all_changes = ("positions", "numbers", "cell", "pbc")

    class PropertyNotImplementedError(NotImplementedError):
        pass

    class _Base:
        def __init__(self, **kwargs):
            self.atoms = None
            self.results = {}

        def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
            if atoms is not None:
                self.atoms = atoms


class SuperlinearCalculator(_Base):
    """ASE calculator over the GPU model (species in model type order)."""

    implemented_properties = ["energy", "free_energy", "forces", "stress"]

    def __init__(self, model, species, precision="fp64", **kwargs):
        super().__init__(**kwargs)
        self.model_path = str(Path(model).expanduser().resolve(strict=True))
        if isinstance(species, str):
            raise ValueError("species must be a sequence of chemical symbols in full model type order")
        self.species = tuple(species)
        if not self.species or not all(isinstance(s, str) for s in self.species):
            raise ValueError("species must contain chemical symbols in full model type order")
        if len(set(self.species)) != len(self.species):
            raise ValueError("species must be nonempty and unique, in exact model type order")
        try:
            numbers = [ATOMIC_NUMBERS[s] for s in self.species]
        except KeyError as exc:
            raise ValueError(f"unknown chemical symbol: {exc.args[0]}") from exc
        self.model = SuperlinearModel(self.model_path, numbers, precision=precision)
        self.calls = 0

    @property
    def metadata(self):
        return {**dict(self.model.metadata), "species": self.species}

    def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
        super().calculate(atoms, properties, system_changes)
        if self.atoms is None:
            raise ValueError("an ASE Atoms object is required")
        atoms = self.atoms
        cell = np.asarray(getattr(atoms.cell, "array", atoms.cell), dtype=np.float64)
        pbc = np.asarray(atoms.pbc, dtype=bool)
        active = cell[pbc]
        if len(active):
            if not np.isfinite(active).all():
                raise ValueError("periodic lattice must be finite")
            if np.linalg.svd(active, compute_uv=False).min() < 0.5:
                raise ValueError("periodic lattice is degenerate or thinner than 0.5 A")
        result = self.model.calculate(np.asarray(atoms.numbers, dtype=np.int32),
                                      np.asarray(atoms.positions, dtype=np.float64), cell, pbc)
        self.calls += 1
        self.results = {
            "energy": float(result["energy"]),
            "free_energy": float(result["energy"]),
            "forces": np.asarray(result["forces"]),
            "superlinear_virial": np.asarray(result["virial"]),
        }
        if result["stress"] is not None:
            self.results["stress"] = np.asarray(result["stress"])
        elif "stress" in properties:
            raise PropertyNotImplementedError("stress requires a finite nonzero 3D cell volume")
