"""Python view of the pair-style compute() contract over the C ABI.

``PairKernel(model, precision)`` mirrors the reference's embedded kernel
(interfaces/lammps/ML-SUS2/sus2_sl_kernel.h:12-43): ``begin(nowned, nall,
types, edges)`` with edges (i owned, j any, d = x_j - x_i), the forward
communication of the gate signal (``pack_forward`` / ``unpack_forward`` over
atom index lists, ``communication_width`` doubles per atom), ``outer()``, the
reverse communication of its adjoint (``pack_reverse`` / ``add_reverse``) and
``finish()`` -> dict(energy, site_energy[nowned], force[nall, 3], virial[6]).
The comm calls take index lists, as ``PairSUS2Embedded::pack_forward_comm``
does (pair_sus2_embedded.cpp:315-319).
"""
import ctypes

import numpy as np

from . import _lib
from .model import PRECISION


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class PairKernel:
    def __init__(self, model, precision="fp64"):
        self.model = model  # Sus2Model (kept alive with the pair handle)
        self._h = ctypes.c_void_p()
        _lib.check(_lib.lib().sus2_pair_create(model.handle, PRECISION[precision], ctypes.byref(self._h)))
        self.nowned = self.nall = 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.lib().sus2_pair_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def communication_width(self):
        return _lib.lib().sus2_pair_comm_width(self._h)

    def begin(self, nowned, nall, types, ei, ej, d):
        types = np.ascontiguousarray(types, np.int32)
        ei = np.ascontiguousarray(ei, np.int32)
        ej = np.ascontiguousarray(ej, np.int32)
        d = np.ascontiguousarray(d, np.float64).reshape(-1, 3)
        _lib.check(_lib.lib().sus2_pair_begin(self._h, int(nowned), int(nall), _p(types), len(ei), _p(ei), _p(ej),
                                              _p(d)))
        self.nowned, self.nall = int(nowned), int(nall)

    def _rows(self, fn, atoms, buf=None):
        atoms = np.ascontiguousarray(atoms, np.int32)
        w = self.communication_width()
        if buf is None:
            buf = np.zeros((len(atoms), w))
        else:
            buf = np.ascontiguousarray(buf, np.float64).reshape(len(atoms), w)
        _lib.check(fn(self._h, len(atoms), _p(atoms), _p(buf)))
        return buf

    def pack_forward(self, atoms):
        return self._rows(_lib.lib().sus2_pair_pack_forward, atoms)

    def unpack_forward(self, atoms, buf):
        self._rows(_lib.lib().sus2_pair_unpack_forward, atoms, buf)

    def outer(self):
        _lib.check(_lib.lib().sus2_pair_outer(self._h))

    def pack_reverse(self, atoms):
        return self._rows(_lib.lib().sus2_pair_pack_reverse, atoms)

    def add_reverse(self, atoms, buf):
        self._rows(_lib.lib().sus2_pair_add_reverse, atoms, buf)

    def finish(self):
        e = np.zeros(1)
        site = np.zeros(max(self.nowned, 1))
        force = np.zeros((max(self.nall, 1), 3))
        vir = np.zeros(6)
        _lib.check(_lib.lib().sus2_pair_finish(self._h, _p(e), _p(site), _p(force), _p(vir)))
        return dict(energy=float(e[0]), site_energy=site[:self.nowned], force=force[:self.nall], virial=vir)
