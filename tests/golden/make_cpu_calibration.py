"""Writes tests/golden/cpu_calibration.json: single-thread CPU time of the
reference's embedded EFS engine (interfaces/lammps/ML-SUS2, built from its
sources into oracle/_ref/ref-timing) against the FP64 oracle's EFS on the same
frames and the same directed edge lists, so that the oracle's throughput on
another host can be restated as a reference-equivalent CPU throughput.

Run in a container that holds the reference tree (make -C oracle ref first):
    python tests/golden/make_cpu_calibration.py
Workloads: the 64-atom Static frame (static64), the 4096-atom Effect frame
(effect4096) and 32 frames of the 4096-frame uneven L4/K6 batch (batch4096,
every 128th frame).  Both sides exclude the neighbour-list build."""
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def write_frames(path, frames):
    with open(path, "w") as fh:
        fh.write(f"{len(frames)}\n")
        for f in frames:
            n = len(f["types"])
            cell = np.asarray(f["cell"], float).ravel()
            fh.write(f"{n}\n" + " ".join(repr(float(c)) for c in cell) + f"\n{1 if f['periodic'] else 0}\n")
            for t, x in zip(f["types"], f["positions"]):
                fh.write(f"{int(t)} {float(x[0])!r} {float(x[1])!r} {float(x[2])!r}\n")


def time_reference(model, frames, budget=6.0):
    binary = oracle.ref_timing_binary()
    with tempfile.TemporaryDirectory() as tmp:
        fp = os.path.join(tmp, "frames.txt")
        write_frames(fp, frames)
        reps = 1
        while True:
            out = subprocess.run([str(binary), model, fp, str(reps)], capture_output=True, text=True, check=True)
            rec = json.loads(out.stdout.split("REF_TIMING_JSON ")[1])
            if rec["seconds_per_rep"] * reps >= budget or reps >= 4096:
                return rec
            reps = max(reps * 2, int(budget / max(rec["seconds_per_rep"], 1e-6)) + 1)


def time_oracle(model, frames, budget=6.0):
    m = oracle.OracleModel(model)
    lists = []
    for f in frames:
        ij, d = oracle.neighbors(f["positions"], f["cell"], f["periodic"], m.cutoff)
        lists.append((f["types"], ij, d))
    reps, energy = 0, 0.0
    t0 = time.perf_counter()
    while True:
        energy = sum(m.efs_edges(*x)[0] for x in lists)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= budget:
            return dict(seconds_per_rep=dt / reps, edges=int(sum(len(x[1]) for x in lists)), energy_sum=energy)


def main():
    from superlinear_amd import workloads as W
    if oracle.ref_timing_binary() is None:
        sys.exit("oracle/_ref/ref-timing missing: run `make -C oracle ref` with the reference tree present")
    cases = {
        "static64": W.config_frames("static64"),
        "effect4096": W.config_frames("effect4096"),
        "batch4096_sample32": W.config_frames("batch4096")[::128],
    }
    out = dict(description="single-thread seconds per EFS evaluation of the reference embedded engine "
                           "(oracle/_ref/ref-timing) and of the FP64 oracle on the same frames and edge lists "
                           "(neighbour build excluded); reference_over_oracle restates oracle throughput as "
                           "reference-equivalent throughput", threads=1,
               host_cpu=open("/proc/cpuinfo").read().split("model name")[1].split(":")[1].split("\n")[0].strip(),
               workloads={})
    for name, frames in cases.items():
        model = W.model_path(W.CONFIG_MODEL[name.split("_")[0]])
        ref = time_reference(model, frames)
        orc = time_oracle(model, frames)
        assert ref["edges"] == orc["edges"], (ref["edges"], orc["edges"])
        assert abs(ref["energy_sum"] - orc["energy_sum"]) <= 1e-9 * (1 + abs(orc["energy_sum"]))
        atoms = sum(len(f["types"]) for f in frames)
        out["workloads"][name] = dict(
            frames=len(frames), atoms=atoms, edges=orc["edges"], model=os.path.basename(model),
            reference_s=ref["seconds_per_rep"], oracle_s=orc["seconds_per_rep"],
            reference_over_oracle=ref["seconds_per_rep"] / orc["seconds_per_rep"],
            reference_atom_steps_per_s=atoms / ref["seconds_per_rep"],
            oracle_atom_steps_per_s=atoms / orc["seconds_per_rep"])
        print(name, out["workloads"][name], flush=True)
    (ROOT / "tests" / "golden" / "cpu_calibration.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    from oracle import oracle
    main()
