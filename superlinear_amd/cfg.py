"""CFG text frames: the data format on both sides of the EFS boundary.

Reader and writer for the reference's text CFG (src/configuration.cpp:429
Load, :1093 Save): ``Size``, optional ``Supercell`` rows (absent = a
non-periodic molecule, src/neighborhoods.cpp:223-227; one or two rows, or a
zero row, leave the remaining directions non-periodic, :490-511), ``AtomData`` with
``cartes_*`` or ``direct_*`` coordinates and optional ``fx fy fz`` and
``site_en`` columns, optional ``Energy`` and ``PlusStress`` (xx yy zz yz xz xy).
"""
import numpy as np

STRESS_ORDER = ("xx", "yy", "zz", "yz", "xz", "xy")


def _frame_from_block(lines):
    f = dict(periodic=False, cell=np.zeros((3, 3)))
    i = 0
    while i < len(lines):
        w = lines[i].split()
        if not w:
            i += 1
            continue
        key = w[0]
        if key == "Size":
            n = int(lines[i + 1].split()[0])
            i += 2
            continue
        if key == "Supercell":
            rows = []
            j = i + 1
            while j < len(lines) and len(rows) < 3:
                p = lines[j].split()
                try:
                    rows.append([float(v) for v in p[:3]])
                except (ValueError, IndexError):
                    break
                j += 1
            if not rows:
                raise ValueError("Supercell without lattice vectors")
            cell = np.zeros((3, 3))  # missing or zero rows: non-periodic directions
            cell[:len(rows)] = rows
            f["cell"], f["periodic"] = cell, bool(np.any(cell != 0))
            i = j
            continue
        if key == "AtomData:":
            cols = w[1:]
            rows = [lines[i + 1 + k].split() for k in range(n)]
            a = np.asarray(rows, dtype=np.float64)
            col = {c: k for k, c in enumerate(cols)}
            f["types"] = a[:, col["type"]].astype(np.int32)
            if "cartes_x" in col:
                f["positions"] = a[:, [col["cartes_x"], col["cartes_y"], col["cartes_z"]]].copy()
            else:
                frac = a[:, [col["direct_x"], col["direct_y"], col["direct_z"]]]
                f["positions"] = frac @ f["cell"]
            if "fx" in col:
                f["forces"] = a[:, [col["fx"], col["fy"], col["fz"]]].copy()
            if "site_en" in col:
                f["site_energy"] = a[:, col["site_en"]].copy()
            i += 1 + n
            continue
        if key == "Energy":
            f["energy"] = float(lines[i + 1].split()[0])
            i += 2
            continue
        if key.startswith("PlusStress"):
            names = lines[i].split(":", 1)[1].split()
            vals = [float(v) for v in lines[i + 1].split()]
            s = dict(zip(names, vals))
            # stored as the 6-vector xx yy zz xy xz yz used by the EFS outputs
            f["virial"] = np.array([s["xx"], s["yy"], s["zz"], s["xy"], s["xz"], s["yz"]])
            i += 2
            continue
        i += 1
    return f


def read_cfg(path):
    """Yield frames (dicts) from a CFG file."""
    block = None
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if s == "BEGIN_CFG":
                block = []
            elif s == "END_CFG":
                if block is not None:
                    yield _frame_from_block(block)
                block = None
            elif block is not None:
                block.append(line)


def write_cfg(fh, frame, lossless=False):
    """Write one frame in the reference's CFG text layout (Save, :1093)."""
    if lossless:
        lat = pos = frc = sen = ene = st = "{:.16e}".format
        ind = typ = "{:d}".format
    else:
        lat = pos = "{:13.6f}".format
        frc = "{:11.6f}".format
        sen = "{:17.12f}".format
        ene = "{:20.12f}".format
        st = "{:11.5f}".format
        ind = "{:10d}".format
        typ = "{:4d}".format
    n = len(frame["types"])
    out = ["BEGIN_CFG", " Size", f"    {n}"]
    if frame.get("periodic", False):
        cell = np.asarray(frame["cell"]).reshape(3, 3)
        count = max([q + 1 for q in range(3) if np.any(cell[q] != 0)], default=0)
        if count:  # rows up to the last non-zero one, as Configuration::Save
            out.append(" Supercell")
            for row in cell[:count]:
                out.append("    " + " ".join(lat(v) for v in row))
    header = " AtomData:  id type       cartes_x      cartes_y      cartes_z"
    if "site_energy" in frame:
        header += "    site_en"
    if "forces" in frame:
        header += "           fx          fy          fz"
    out.append(header)
    x = np.asarray(frame["positions"]).reshape(n, 3)
    for i in range(n):
        row = f"    {ind(i + 1)} {typ(int(frame['types'][i]))}  " + " ".join(pos(v) for v in x[i])
        if "site_energy" in frame:
            row += "  " + sen(frame["site_energy"][i])
        if "forces" in frame:
            row += "  " + " ".join(frc(v) for v in np.asarray(frame["forces"])[i])
        out.append(row)
    if "energy" in frame:
        out += [" Energy", "    " + ene(frame["energy"])]
    if "virial" in frame:
        w = frame["virial"]  # xx yy zz xy xz yz
        out.append(" PlusStress:  xx          yy          zz          yz          xz          xy")
        out.append("     " + " ".join(st(v) for v in (w[0], w[1], w[2], w[5], w[4], w[3])))
    out.append("END_CFG")
    fh.write("\n".join(out) + "\n\n")
