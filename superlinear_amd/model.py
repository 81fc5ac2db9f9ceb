"""Host mirror of the reference model operator (MLMTPR / AnyLocalMLIP) over
the MI355X C ABI.

``Sus2Model(path)`` loads a SUS2-SH text model; ``calc_efs`` evaluates a batch
of frames on the GPU (``precision`` "fp64" or "fp32") and returns energies,
forces, PlusStress virials and site energies, as ``MLMTPR::CalcEFS`` does per
frame (dev_src/mtpr.cpp:8779).  ``coeffs`` / ``set_coeffs`` expose the flat
coefficient vector in the reference optimiser layout (dev_src/mtpr.cpp
CoeffCount/Coeff).
"""
import ctypes

import numpy as np

from . import _lib

PRECISION = {"fp64": _lib.SUS2_FP64, "fp32": _lib.SUS2_FP32, 0: 0, 1: 1}
DIM_NAMES = ("C", "L", "K", "B", "F", "M", "Q", "R", "O", "Qg", "Kw", "nbasic", "nprod", "gated",
             "species_readout", "H")
OFF_NAMES = ("shift", "scal", "rad", "grad", "gscal", "alpha", "gtype", "W", "gw", "ow", "species", "read", "ncoeff")


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def pack_frames(frames):
    """Concatenate frames into the batch arrays of the C ABI."""
    off = np.zeros(len(frames) + 1, np.int64)
    for k, f in enumerate(frames):
        off[k + 1] = off[k] + len(f["types"])
    types = np.ascontiguousarray(np.concatenate([np.asarray(f["types"], np.int32) for f in frames])
                                 if frames else np.zeros(0, np.int32), np.int32)
    pos = np.ascontiguousarray(np.concatenate([np.asarray(f["positions"], np.float64).reshape(-1, 3)
                                               for f in frames]) if frames else np.zeros((0, 3)), np.float64)
    cell = np.ascontiguousarray(np.stack([np.asarray(f.get("cell", np.zeros((3, 3))), np.float64).reshape(3, 3)
                                          for f in frames]) if frames else np.zeros((0, 3, 3)), np.float64)
    periodic = np.asarray([1 if f.get("periodic", True) else 0 for f in frames], np.int32)
    return dict(off=off, types=types, pos=pos, cell=cell, periodic=periodic, nframes=len(frames),
                natoms=int(off[-1]))


class Sus2Model:
    def __init__(self, path):
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().sus2_model_load(str(path).encode(), ctypes.byref(self._h)))
        self.path = str(path)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().sus2_model_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def cutoff(self):
        return _lib.lib().sus2_model_cutoff(self._h)

    def dims(self):
        d = np.zeros(16, np.int32)
        _lib.check(_lib.lib().sus2_model_dims(self._h, d.ctypes.data_as(_lib._ip)))
        return dict(zip(DIM_NAMES, d.tolist()))

    def offsets(self):
        d = np.zeros(13, np.int32)
        _lib.check(_lib.lib().sus2_coeff_offsets(self._h, d.ctypes.data_as(_lib._ip)))
        return dict(zip(OFF_NAMES, d.tolist()))

    @property
    def coeff_count(self):
        return _lib.lib().sus2_coeff_count(self._h)

    def coeffs(self):
        out = np.zeros(self.coeff_count)
        _lib.check(_lib.lib().sus2_coeff_get(self._h, out.ctypes.data_as(_lib._dp)))
        return out

    def set_coeffs(self, theta):
        theta = np.ascontiguousarray(theta, np.float64)
        if theta.size != self.coeff_count:
            raise ValueError(f"expected {self.coeff_count} coefficients, got {theta.size}")
        _lib.check(_lib.lib().sus2_coeff_set(self._h, theta.ctypes.data_as(_lib._dp)))

    def save(self, path):
        _lib.check(_lib.lib().sus2_model_save(self._h, str(path).encode()))

    # ------------------------------------------------------------ diagnostics
    def last_edge_count(self):
        n = ctypes.c_int64()
        _lib.check(_lib.lib().sus2_last_edge_count(self._h, ctypes.byref(n)))
        return n.value

    def profile(self, on=True):
        _lib.check(_lib.lib().sus2_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self):
        _lib.check(_lib.lib().sus2_profile_reset(self._h))

    def profile_read(self):
        """{kernel name: (total ms, launches)} from the HIP-event timers."""
        cap = 64
        names = ctypes.create_string_buffer(64 * cap)
        ms = np.zeros(cap)
        cnt = np.zeros(cap, np.int64)
        n = ctypes.c_int()
        _lib.check(_lib.lib().sus2_profile_read(self._h, cap, names, ms.ctypes.data_as(_lib._dp),
                                                cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                ctypes.byref(n)))
        raw = names.raw
        out = {}
        for k in range(min(n.value, cap)):
            nm = raw[64 * k:64 * (k + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[k]), int(cnt[k]))
        return out

    # ------------------------------------------------------------ evaluation
    def calc_efs(self, frames, precision="fp64"):
        """EFS of a list of frames (host arrays in, host arrays out)."""
        b = pack_frames(frames)
        nf, n = b["nframes"], b["natoms"]
        E = np.zeros(nf)
        F = np.zeros((n, 3))
        W = np.zeros((nf, 6))
        S = np.zeros(n)
        _lib.check(_lib.lib().sus2_efs_host(self._h, PRECISION[precision], nf, n, _p(b["off"]), _p(b["types"]),
                                            _p(b["pos"]), _p(b["cell"]), _p(b["periodic"]), _p(E), _p(F), _p(W),
                                            _p(S)))
        return dict(energy=E, forces=F, virial=W, site=S)

    def efs_device(self, batch, precision="fp64", out=None, stream=None):
        """EFS on device-resident torch tensors (see DeviceBatch); returns the
        output tensors (energy[nf], forces[n,3], virial[nf,6], site[n])."""
        import torch
        nf, n = batch.nframes, batch.natoms
        dev = batch.pos.device
        if out is None:
            out = dict(energy=torch.empty(nf, dtype=torch.float64, device=dev),
                       forces=torch.empty((n, 3), dtype=torch.float64, device=dev),
                       virial=torch.empty((nf, 6), dtype=torch.float64, device=dev),
                       site=torch.empty(n, dtype=torch.float64, device=dev))
        st = stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream
        _lib.check(_lib.lib().sus2_efs(self._h, PRECISION[precision], nf, n, batch.off.data_ptr(),
                                       batch.types.data_ptr(), batch.pos.data_ptr(), batch.cell.data_ptr(),
                                       batch.periodic.data_ptr(), out["energy"].data_ptr(),
                                       out["forces"].data_ptr(), out["virial"].data_ptr(),
                                       out["site"].data_ptr() if out.get("site") is not None else None,
                                       ctypes.c_void_p(st)))
        return out


class DeviceBatch:
    """Frames packed into device-resident torch tensors (HBM), the layout the
    device entry points read."""

    def __init__(self, frames, device="cuda"):
        import torch
        b = pack_frames(frames)
        self.nframes, self.natoms = b["nframes"], b["natoms"]
        self.off = torch.from_numpy(b["off"]).to(device)
        self.types = torch.from_numpy(b["types"]).to(device)
        self.pos = torch.from_numpy(b["pos"]).to(device)
        self.cell = torch.from_numpy(b["cell"]).to(device)
        self.periodic = torch.from_numpy(b["periodic"]).to(device)
        self.host = b
