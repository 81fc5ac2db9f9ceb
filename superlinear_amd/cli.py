"""Command-line drop-ins for the reference's EFS verbs, evaluated on the GPU.

    python -m superlinear_amd.cli calc-efs pot.sus2 in.cfg out.cfg [--no-loss]
    python -m superlinear_amd.cli calc-errors pot.sus2 db.cfg

``calc-efs`` (dev_src/mlp/dev_mlp_commands.cpp:3037-3068) writes every
configuration of in.cfg with its energy, forces, PlusStress and site energies
(the ``site_en`` column holds the readout part of the site energy, the
reference's cal_se0, dev_src/mtpr_sh.cpp:10572-10593).  ``--no-loss`` keeps
full double precision.  ``calc-errors`` (src/mlp/mlp_commands.cpp:793-808, the
metrics of dev_src/non_linear_regression.cpp:779-804 and the summary block of
dev_src/mlp/mtpr_train.cpp:3893-3903) prints energy / force / stress MAE and
RMSE of the model against the labels in db.cfg.  Frames are evaluated in
batches of up to ``--batch-atoms`` atoms in one GPU call; ``--precision
fp32`` selects the FP32 kernels.
"""
import argparse
import sys

import numpy as np

from .cfg import read_cfg, write_cfg
from .model import Sus2Model


def _batches(frames, max_atoms):
    batch, n = [], 0
    for f in frames:
        k = len(f["types"])
        if batch and n + k > max_atoms:
            yield batch
            batch, n = [], 0
        batch.append(f)
        n += k
    if batch:
        yield batch


def _evaluate(model, frames, precision, max_atoms):
    """Yield (frame, energy, forces, virial, site) in input order."""
    for batch in _batches(frames, max_atoms):
        r = model.calc_efs(batch, precision)
        a = 0
        for k, f in enumerate(batch):
            n = len(f["types"])
            yield f, r["energy"][k], r["forces"][a:a + n], r["virial"][k], r["site"][a:a + n]
            a += n


def calc_efs(pot, inp, out, lossless=False, precision="fp64", max_atoms=1 << 20):
    model = Sus2Model(pot)
    off = model.offsets()
    theta = model.coeffs()
    const = theta[off["species"]:off["species"] + model.dims()["C"]] + theta[off["shift"]:off["shift"] +
                                                                               model.dims()["C"]]
    count = 0
    with open(out, "w") as fh:
        for f, e, frc, vir, site in _evaluate(model, read_cfg(inp), precision, max_atoms):
            g = dict(f)
            g["energy"] = float(e)
            g["forces"] = frc
            g["virial"] = vir
            g["site_energy"] = site - const[np.asarray(f["types"])]
            write_cfg(fh, g, lossless=lossless)
            count += 1
    return count


def calc_errors(pot, db, precision="fp64", max_atoms=1 << 20, stream=sys.stdout):
    model = Sus2Model(pot)
    cfgs = atoms = 0
    e_abs = e_sq = 0.0
    f_abs = f_sq = 0.0
    f_cnt = 0
    s_abs = s_sq = 0.0
    s_cnt = 0
    for f, e, frc, vir, _ in _evaluate(model, read_cfg(db), precision, max_atoms):
        n = len(f["types"])
        cfgs += 1
        if "energy" in f:
            d = f["energy"] - e
            e_abs += abs(d)
            e_sq += d * d / n
            atoms += n
        if "forces" in f:
            d = frc - np.asarray(f["forces"])
            f_abs += float(np.abs(d).sum())
            f_sq += float((d * d).sum())
            f_cnt += 3 * n
        if "virial" in f:
            d = np.asarray(vir) - np.asarray(f["virial"])
            s_abs += float(np.abs(d).sum())
            s_sq += float((d * d).sum())
            s_cnt += 6
    m = dict(
        structures=cfgs,
        atoms=atoms,
        energy_mae_mev_atom=1000.0 * e_abs / atoms if atoms else 0.0,
        energy_rmse_mev_atom=1000.0 * np.sqrt(e_sq / atoms) if atoms else 0.0,
        force_mae_mev_a=1000.0 * f_abs / f_cnt if f_cnt else 0.0,
        force_rmse_mev_a=1000.0 * np.sqrt(f_sq / f_cnt) if f_cnt else 0.0,
        stress_mae_ev=s_abs / s_cnt if s_cnt else 0.0,
        stress_rmse_ev=np.sqrt(s_sq / s_cnt) if s_cnt else 0.0,
    )
    if stream is not None:
        stream.write("\n=== Train Summary ===\n"
                     f"Structures           : {m['structures']}\n"
                     f"Atoms                : {m['atoms']}\n"
                     f"Energy MAE (meV/atom): {m['energy_mae_mev_atom']:.3f}\n"
                     f"Energy RMSE(meV/atom): {m['energy_rmse_mev_atom']:.3f}\n"
                     f"Force MAE  (meV/A)   : {m['force_mae_mev_a']:.3f}\n"
                     f"Force RMSE (meV/A)   : {m['force_rmse_mev_a']:.3f}\n"
                     f"Stress MAE (eV)      : {m['stress_mae_ev']:.3f}\n"
                     f"Stress RMSE(eV)      : {m['stress_rmse_ev']:.3f}\n"
                     "=====================\n")
    return m


def main(argv=None):
    ap = argparse.ArgumentParser(prog="superlinear_amd.cli")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("calc-efs", help="energies, forces and stresses of configurations")
    a.add_argument("pot")
    a.add_argument("inp")
    a.add_argument("out")
    a.add_argument("--no-loss", action="store_true", help="preserve full double precision in the output CFG")
    b = sub.add_parser("calc-errors", help="errors of the potential on a labelled database")
    b.add_argument("pot")
    b.add_argument("db")
    for p in (a, b):
        p.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
        p.add_argument("--batch-atoms", type=int, default=1 << 20)
    args = ap.parse_args(argv)
    if args.cmd == "calc-efs":
        calc_efs(args.pot, args.inp, args.out, args.no_loss, args.precision, args.batch_atoms)
    else:
        calc_errors(args.pot, args.db, args.precision, args.batch_atoms)
    return 0


if __name__ == "__main__":
    sys.exit(main())
