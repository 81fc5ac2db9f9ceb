"""ctypes view of the FP64 oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, always as the checker.  Nothing under superlinear_amd/ imports it.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def _ptr(a, kind):
    return a.ctypes.data_as(kind) if a is not None else None


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(str(path))
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_model_load.restype = ctypes.c_void_p
        L.orc_model_load.argtypes = [ctypes.c_char_p]
        L.orc_model_free.argtypes = [ctypes.c_void_p]
        L.orc_coeff_count.argtypes = [ctypes.c_void_p]
        L.orc_coeff_get.argtypes = [ctypes.c_void_p, _dp]
        L.orc_coeff_set.argtypes = [ctypes.c_void_p, _dp]
        L.orc_cutoff.argtypes = [ctypes.c_void_p]
        L.orc_cutoff.restype = ctypes.c_double
        L.orc_dims.argtypes = [ctypes.c_void_p, _ip]
        L.orc_offsets.argtypes = [ctypes.c_void_p, _ip]
        L.orc_sh.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_neighbors.restype = ctypes.c_long
        L.orc_neighbors.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_long, _ip, _dp]
        L.orc_efs_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _ip, _dp, _dp, _ip, _dp, _dp, _dp, _dp,
                                    ctypes.c_int]
        L.orc_efs_edges.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_long, _ip, _dp,
                                    _dp, _dp, _dp, _dp]
        L.orc_loss_grad.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _ip, _dp, _dp, _ip, _dp, _dp, _dp,
                                    ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                                    _dp, _dp, _dp, ctypes.c_int]
        L.orc_efs_grad.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _ip, _dp, _dp, _ip, _dp, _dp, _dp, _dp,
                                   ctypes.c_int]
        L.orc_loss.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _ip, _dp, _dp, _ip, _dp, _dp, _dp,
                               ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_set_grad_subset.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.orc_frame_loss.restype = ctypes.c_double
        L.orc_frame_loss.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp, _ip, _dp, _dp,
                                     _dp]
        L.orc_radial_penalty.restype = ctypes.c_double
        L.orc_radial_penalty.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


DIM_NAMES = ("C", "L", "K", "B", "F", "M", "Q", "R", "O", "Qg", "Kw", "nbasic", "nprod")
OFF_NAMES = ("shift", "scal", "rad", "grad", "gscal", "alpha", "gtype", "W", "gw", "ow", "species", "read", "ncoeff")


class OracleModel:
    """FP64 oracle model handle (own parser of the SUS2-SH text format)."""

    def __init__(self, path):
        self.h = lib().orc_model_load(str(path).encode())
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_model_free(self.h)
            self.h = None

    @property
    def cutoff(self):
        return lib().orc_cutoff(self.h)

    def dims(self):
        d = np.zeros(13, np.int32)
        lib().orc_dims(self.h, _ptr(d, _ip))
        return dict(zip(DIM_NAMES, d.tolist()))

    def offsets(self):
        d = np.zeros(13, np.int32)
        lib().orc_offsets(self.h, _ptr(d, _ip))
        return dict(zip(OFF_NAMES, d.tolist()))

    def coeffs(self):
        n = lib().orc_coeff_count(self.h)
        out = np.zeros(n)
        lib().orc_coeff_get(self.h, _ptr(out, _dp))
        return out

    def set_coeffs(self, theta):
        theta = np.ascontiguousarray(theta, np.float64)
        assert theta.size == lib().orc_coeff_count(self.h)
        lib().orc_coeff_set(self.h, _ptr(theta, _dp))

    def efs(self, frames, nthreads=1):
        """frames: list of dicts(types, positions, cell, periodic). Returns
        dict(energy[nf], forces[N,3], virial[nf,6], site[N])."""
        b = pack(frames)
        nf = len(frames)
        n = int(b["off"][-1])
        E = np.zeros(nf)
        F = np.zeros((n, 3))
        W = np.zeros((nf, 6))
        S = np.zeros(n)
        _check(lib().orc_efs_batch(self.h, nf, _ptr(b["off"], _lp), _ptr(b["types"], _ip), _ptr(b["pos"], _dp),
                                   _ptr(b["cell"], _dp), _ptr(b["periodic"], _ip), _ptr(E, _dp), _ptr(F, _dp),
                                   _ptr(W, _dp), _ptr(S, _dp), nthreads))
        return dict(energy=E, forces=F, virial=W, site=S)

    def grad_subset(self, idx=None):
        """Compute only the listed gradient entries (None: all); the others
        come back as zero."""
        a = np.ascontiguousarray(np.zeros(0) if idx is None else idx, np.int32)
        _check(lib().orc_set_grad_subset(self.h, a.size, _ptr(a, _ip)))

    def radial_penalty(self, constr, theta=None):
        """The legacy radial penalty of theta (default: the model's) and its gradient."""
        th = np.ascontiguousarray(self.coeffs() if theta is None else theta, np.float64)
        g = np.zeros(th.size)
        v = lib().orc_radial_penalty(self.h, _ptr(th, _dp), float(constr), _ptr(g, _dp))
        return v, g

    def loss_grad(self, frames, labels, mode=1, wE=1.0, wF=0.01, wS=0.0, nthreads=1, objective=None):
        b = pack(frames)
        nf = len(frames)
        n = int(b["off"][-1])
        E0 = np.ascontiguousarray(labels["energy"], np.float64)
        F0 = np.ascontiguousarray(labels["forces"], np.float64).reshape(n, 3)
        S0 = np.ascontiguousarray(labels.get("virial", np.zeros((nf, 6))), np.float64).reshape(nf, 6)
        loss = np.zeros(1)
        grad = np.zeros(lib().orc_coeff_count(self.h))
        eb = np.zeros(n)
        fb = np.zeros((n, 3))
        sb = np.zeros((nf, 9))
        _check(lib().orc_loss_grad(self.h, nf, _ptr(b["off"], _lp), _ptr(b["types"], _ip), _ptr(b["pos"], _dp),
                                   _ptr(b["cell"], _dp), _ptr(b["periodic"], _ip), _ptr(E0, _dp), _ptr(F0, _dp),
                                   _ptr(S0, _dp), mode, wE, wF, wS, _ptr(objective_array(objective), _dp),
                                   _ptr(loss, _dp), _ptr(grad, _dp), _ptr(eb, _dp), _ptr(fb, _dp), _ptr(sb, _dp),
                                   nthreads))
        return dict(loss=float(loss[0]), grad=grad, ebar=eb, fbar=fb, sbar=sb)

    def efs_grad(self, frames, ene_w, frc_w, str_w, nthreads=1):
        """sum over frames of sum_i ene_w_i dE_i/dtheta + d/dtheta (sum_e w_e . dE/dd_e),
        w_e = frc_w_i - frc_w_j + str_w(frame) d_e (str_w row-major 3x3 per frame)."""
        b = pack(frames)
        nf = len(frames)
        n = int(b["off"][-1])
        ew = np.ascontiguousarray(ene_w, np.float64).reshape(n)
        fw = np.ascontiguousarray(frc_w, np.float64).reshape(n, 3)
        sw = np.ascontiguousarray(str_w, np.float64).reshape(nf, 9)
        grad = np.zeros(lib().orc_coeff_count(self.h))
        _check(lib().orc_efs_grad(self.h, nf, _ptr(b["off"], _lp), _ptr(b["types"], _ip), _ptr(b["pos"], _dp),
                                  _ptr(b["cell"], _dp), _ptr(b["periodic"], _ip), _ptr(ew, _dp), _ptr(fw, _dp),
                                  _ptr(sw, _dp), _ptr(grad, _dp), nthreads))
        return grad

    def efs_edges(self, types, ij, d):
        """EFS of one domain owning every atom from a given directed edge list
        (ij [E,2], d [E,3] = x_j - x_i): returns (energy, forces [n,3])."""
        types = np.ascontiguousarray(types, np.int32)
        ij = np.ascontiguousarray(ij, np.int32)
        d = np.ascontiguousarray(d, np.float64)
        n = len(types)
        E = np.zeros(1)
        F = np.zeros((n, 3))
        W = np.zeros(6)
        S = np.zeros(n)
        _check(lib().orc_efs_edges(self.h, n, n, _ptr(types, _ip), len(ij), _ptr(ij, _ip), _ptr(d, _dp),
                                   _ptr(E, _dp), _ptr(F, _dp), _ptr(W, _dp), _ptr(S, _dp)))
        return float(E[0]), F

    def loss(self, frames, labels, mode=1, wE=1.0, wF=0.01, wS=0.0, objective=None):
        b = pack(frames)
        nf = len(frames)
        n = int(b["off"][-1])
        E0 = np.ascontiguousarray(labels["energy"], np.float64)
        F0 = np.ascontiguousarray(labels["forces"], np.float64).reshape(n, 3)
        S0 = np.ascontiguousarray(labels.get("virial", np.zeros((nf, 6))), np.float64).reshape(nf, 6)
        loss = np.zeros(1)
        _check(lib().orc_loss(self.h, nf, _ptr(b["off"], _lp), _ptr(b["types"], _ip), _ptr(b["pos"], _dp),
                              _ptr(b["cell"], _dp), _ptr(b["periodic"], _ip), _ptr(E0, _dp), _ptr(F0, _dp),
                              _ptr(S0, _dp), mode, wE, wF, wS, _ptr(objective_array(objective), _dp),
                              _ptr(loss, _dp)))
        return float(loss[0])


# objective options in the order of the C functions' opts[7]; defaults are
# those of `mlp-sus2 train` for Static and Effect models
OBJECTIVE_KEYS = ("rel_forces", "norm_by_forces", "force_loss", "log_cosh_scale", "std_weight", "stdd_weight",
                  "constr_weight")
OBJECTIVE_DEFAULT = dict(rel_forces=0.0, norm_by_forces=0, force_loss="l2", log_cosh_scale=2.0, std_weight=0.0,
                         stdd_weight=0.0, constr_weight=1e-9)


def objective_array(objective=None):
    o = dict(OBJECTIVE_DEFAULT)
    if objective:
        unknown = set(objective) - set(o)
        if unknown:
            raise ValueError(f"unknown objective options {sorted(unknown)}")
        o.update(objective)
    o["force_loss"] = {"l2": 0, "log-cosh": 1}.get(o["force_loss"], o["force_loss"])
    return np.array([float(o[k]) for k in OBJECTIVE_KEYS], np.float64)


def frame_loss(mode, wE, wF, wS, E, F, S, E0, F0, S0, objective=None, site=None, types=None):
    """One frame's loss and adjoints from given predictions (E, F [n,3], S
    Voigt [6]) and labels, under the objective options (site energies and
    species only for the std terms): returns (loss, dL/dE_i [n], dL/dF [n,3],
    dL/dS [3,3])."""
    F = np.ascontiguousarray(F, np.float64).reshape(-1, 3)
    F0 = np.ascontiguousarray(F0, np.float64).reshape(-1, 3)
    S = np.ascontiguousarray(S, np.float64)
    S0 = np.ascontiguousarray(S0, np.float64)
    n = len(F)
    dE = np.zeros(n)
    dF = np.zeros((n, 3))
    dS = np.zeros(9)
    o = objective_array(objective)
    se = None if site is None else np.ascontiguousarray(site, np.float64)
    ty = None if types is None else np.ascontiguousarray(types, np.int32)
    loss = lib().orc_frame_loss(n, mode, wE, wF, wS, E, _ptr(F, _dp), _ptr(S, _dp), E0, _ptr(F0, _dp),
                                _ptr(S0, _dp), _ptr(o, _dp), None if se is None else _ptr(se, _dp),
                                None if ty is None else _ptr(ty, _ip), _ptr(dE, _dp), _ptr(dF, _dp), _ptr(dS, _dp))
    if np.isnan(loss):
        raise RuntimeError(lib().orc_last_error().decode())
    return loss, dE, dF, dS.reshape(3, 3)


def pack(frames):
    off = np.zeros(len(frames) + 1, np.int64)
    for k, f in enumerate(frames):
        off[k + 1] = off[k] + len(f["types"])
    types = np.ascontiguousarray(np.concatenate([np.asarray(f["types"], np.int32) for f in frames]), np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(f["positions"], np.float64).reshape(-1, 3)
                                               for f in frames]), np.float64)
    cell = np.ascontiguousarray(np.stack([np.asarray(f.get("cell", np.zeros((3, 3))), np.float64).reshape(3, 3)
                                          for f in frames]), np.float64)
    periodic = np.asarray([1 if f.get("periodic", True) else 0 for f in frames], np.int32)
    return dict(off=off, types=types, pos=pos, cell=cell, periodic=periodic)


def sh(L, d):
    d = np.ascontiguousarray(d, np.float64)
    H = (L + 1) ** 2
    Y = np.zeros(H)
    dY = np.zeros((H, 3))
    lib().orc_sh(L, _ptr(d, _dp), _ptr(Y, _dp), _ptr(dY, _dp))
    return Y, dY


def neighbors(positions, cell, periodic, rc):
    x = np.ascontiguousarray(positions, np.float64)
    c = np.ascontiguousarray(cell, np.float64)
    n = len(x)
    cnt = lib().orc_neighbors(n, _ptr(x, _dp), _ptr(c, _dp), int(periodic), rc, 0, None, None)
    if cnt < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    ij = np.zeros((cnt, 2), np.int32)
    d = np.zeros((cnt, 3))
    lib().orc_neighbors(n, _ptr(x, _dp), _ptr(c, _dp), int(periodic), rc, cnt, _ptr(ij, _ip), _ptr(d, _dp))
    return ij, d


def build():
    """Compile liboracle.so (and the reference engine when its tree exists)."""
    import subprocess
    subprocess.run(["make", "-C", str(_HERE), "liboracle.so"], check=True, stdout=subprocess.DEVNULL)
    if Path("/root/reference/interfaces/lammps/ML-SUS2").is_dir():
        subprocess.run(["make", "-C", str(_HERE), "ref"], check=True, stdout=subprocess.DEVNULL)


def ref_timing_binary():
    p = _HERE / "_ref" / "ref-timing"
    return p if p.exists() and os.access(p, os.X_OK) else None


def ref_binary():
    p = _HERE / "_ref" / "embedded-efs"
    return p if p.exists() and os.access(p, os.X_OK) else None
