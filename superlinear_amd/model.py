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

PRECISION = {"fp64": _lib.SUS2_FP64, "fp32": _lib.SUS2_FP32, "mixed": _lib.SUS2_MIXED, 0: 0, 1: 1, 2: 2}
FORCE_LOSS = {"l2": 0, "log-cosh": 1}
FORCE_LOSS_NAME = {0: "l2", 1: "log-cosh"}
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


def _stream_of(stream, dev):
    """The HIP stream handle a call runs on: the caller's, else torch's current."""
    import torch
    return stream if stream is not None else torch.cuda.current_stream(dev).cuda_stream


def _join(stream, dev):
    """After a call on a caller-given stream: torch's current stream waits for
    it, so that the returned tensors (allocated on the current stream) are not
    read or recycled before the kernels writing them have finished."""
    import torch
    if stream is None:
        return
    ext = torch.cuda.ExternalStream(stream, device=dev)
    cur = torch.cuda.current_stream(dev)
    if ext.cuda_stream != cur.cuda_stream:
        cur.wait_stream(ext)


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

    # ------------------------------------------------------------ objective options
    def objective(self):
        """The handle's objective options (sus2_objective) as a dict; a new
        handle holds the options of `mlp-sus2 train` for these models."""
        o = _lib.Sus2Objective()
        _lib.check(_lib.lib().sus2_objective_get(self._h, ctypes.byref(o)))
        d = {name: getattr(o, name) for name, _ in o._fields_}
        d["force_loss"] = FORCE_LOSS_NAME[d["force_loss"]]
        return d

    def set_objective(self, **opts):
        """Change objective options (rel_forces, norm_by_forces, force_loss
        'l2' / 'log-cosh', log_cosh_scale, std_weight, stdd_weight,
        constr_weight); unnamed options keep their values.  Used by the
        objective calls, trainers and graphs made afterwards."""
        cur = self.objective()
        unknown = set(opts) - set(cur)
        if unknown:
            raise ValueError(f"unknown objective options {sorted(unknown)}")
        cur.update(opts)
        cur["force_loss"] = FORCE_LOSS.get(cur["force_loss"], cur["force_loss"])
        o = _lib.Sus2Objective(**cur)
        _lib.check(_lib.lib().sus2_objective_set(self._h, ctypes.byref(o)))

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
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_efs(self._h, PRECISION[precision], nf, n, batch.off.data_ptr(),
                                       batch.types.data_ptr(), batch.pos.data_ptr(), batch.cell.data_ptr(),
                                       batch.periodic.data_ptr(), out["energy"].data_ptr(),
                                       out["forces"].data_ptr(), out["virial"].data_ptr(),
                                       out["site"].data_ptr() if out.get("site") is not None else None,
                                       ctypes.c_void_p(st)))
        _join(stream, dev)
        return out


    # ------------------------------------------------------------ gradients
    def efs_grad_device(self, batch, ene_w, frc_w, str_w, precision="fp64", grad=None, stream=None):
        """grad += d/dtheta [sum_i ene_w_i E_i + sum_e w_e . dE/dd_e] with
        w_e = frc_w_i - frc_w_j + str_w(frame) d_e  (AccumulateEFSCombinationGrad,
        dev_src/mtpr.cpp:9188).  Weights are FP64 device tensors: ene_w [N],
        frc_w [N,3], str_w [nf,3,3]; grad is an FP64 device tensor [ncoeff]."""
        import torch
        dev = batch.pos.device
        if grad is None:
            grad = torch.zeros(self.coeff_count, dtype=torch.float64, device=dev)
        ew = ene_w.to(device=dev, dtype=torch.float64).contiguous()
        fw = frc_w.to(device=dev, dtype=torch.float64).contiguous()
        sw = str_w.to(device=dev, dtype=torch.float64).contiguous()
        if ew.numel() != batch.natoms or fw.numel() != 3 * batch.natoms or sw.numel() != 9 * batch.nframes:
            raise ValueError("weight shapes do not match the batch")
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_efs_grad(self._h, PRECISION[precision], batch.nframes, batch.natoms,
                                            batch.off.data_ptr(), batch.types.data_ptr(), batch.pos.data_ptr(),
                                            batch.cell.data_ptr(), batch.periodic.data_ptr(), ew.data_ptr(),
                                            fw.data_ptr(), sw.data_ptr(), grad.data_ptr(), ctypes.c_void_p(st)))
        _join(stream, dev)
        return grad

    def loss_grad_device(self, batch, e0, f0, s0=None, weighting="structures", wE=1.0, wF=0.01, wS=0.0,
                         precision="fp32", grad=None, stream=None):
        """Training objective (1/nframes) sum_f L_f + penalty (the handle's
        objective options) and its gradient
        (NonLinearRegression::CalcObjectiveFunctionGrad,
        dev_src/non_linear_regression.cpp:1112).  Labels are device tensors:
        e0 [nf], f0 [N,3], s0 [nf,6] (xx yy zz xy xz yz).  Returns
        (loss, grad) with grad an FP64 device tensor (overwritten).  The call
        returns once the loss is known; the gradient kernels finish on the
        call's stream, which torch's current stream is made to wait for."""
        import torch
        dev = batch.pos.device
        mode = {"vibrations": 0, "structures": 1, "molecules": 2}.get(weighting, weighting)
        if grad is None:
            grad = torch.empty(self.coeff_count, dtype=torch.float64, device=dev)
        e0 = e0.to(device=dev, dtype=torch.float64).contiguous()
        f0 = f0.to(device=dev, dtype=torch.float64).contiguous()
        if s0 is None:  # zero stress labels, kept between calls (only read)
            z = getattr(self, "_s0_zero", None)
            if z is None or z.device != dev or z.shape[0] != batch.nframes:
                z = self._s0_zero = torch.zeros((batch.nframes, 6), dtype=torch.float64, device=dev)
            s0 = z
        else:
            s0 = s0.to(device=dev, dtype=torch.float64).contiguous()
        loss = ctypes.c_double()
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_loss_grad(self._h, PRECISION[precision], batch.nframes, batch.natoms,
                                             batch.off.data_ptr(), batch.types.data_ptr(), batch.pos.data_ptr(),
                                             batch.cell.data_ptr(), batch.periodic.data_ptr(), e0.data_ptr(),
                                             f0.data_ptr(), s0.data_ptr(), int(mode), float(wE), float(wF),
                                             float(wS), ctypes.byref(loss), grad.data_ptr(), ctypes.c_void_p(st)))
        _join(stream, dev)
        return loss.value, grad

    # ------------------------------------------------------------ graphs and launch shape
    def graph_efs(self, batch, precision="fp32", stream=None):
        """Capture the batch EFS in a HIP graph (outputs allocated here)."""
        import torch
        nf, n = batch.nframes, batch.natoms
        dev = batch.pos.device
        out = dict(energy=torch.zeros(nf, dtype=torch.float64, device=dev),
                   forces=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                   virial=torch.zeros((nf, 6), dtype=torch.float64, device=dev),
                   site=torch.zeros(n, dtype=torch.float64, device=dev))
        h = ctypes.c_void_p()
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_graph_efs_create(
            self._h, PRECISION[precision], nf, n, batch.off.data_ptr(), batch.types.data_ptr(),
            batch.pos.data_ptr(), batch.cell.data_ptr(), batch.periodic.data_ptr(), out["energy"].data_ptr(),
            out["forces"].data_ptr(), out["virial"].data_ptr(), out["site"].data_ptr(), ctypes.c_void_p(st),
            ctypes.byref(h)))
        return Sus2Graph(h, [batch], out)

    def graph_loss_grad(self, batch, e0, f0, s0=None, weighting="structures", wE=1.0, wF=0.01, wS=0.0,
                        precision="fp32", stream=None):
        """Capture the training objective + gradient (and the batch's EFS) in a
        HIP graph; outputs: loss [1], grad [ncoeff], energy, forces, virial."""
        import torch
        dev = batch.pos.device
        nf, n = batch.nframes, batch.natoms
        mode = {"vibrations": 0, "structures": 1, "molecules": 2}.get(weighting, weighting)
        e0 = e0.to(device=dev, dtype=torch.float64).contiguous()
        f0 = f0.to(device=dev, dtype=torch.float64).contiguous()
        s0 = (torch.zeros((nf, 6), dtype=torch.float64, device=dev) if s0 is None
              else s0.to(device=dev, dtype=torch.float64).contiguous())
        out = dict(loss=torch.zeros(1, dtype=torch.float64, device=dev),
                   grad=torch.zeros(self.coeff_count, dtype=torch.float64, device=dev),
                   energy=torch.zeros(nf, dtype=torch.float64, device=dev),
                   forces=torch.zeros((n, 3), dtype=torch.float64, device=dev),
                   virial=torch.zeros((nf, 6), dtype=torch.float64, device=dev))
        h = ctypes.c_void_p()
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_graph_loss_grad_create(
            self._h, PRECISION[precision], nf, n, batch.off.data_ptr(), batch.types.data_ptr(),
            batch.pos.data_ptr(), batch.cell.data_ptr(), batch.periodic.data_ptr(), e0.data_ptr(), f0.data_ptr(),
            s0.data_ptr(), int(mode), float(wE), float(wF), float(wS), out["loss"].data_ptr(),
            out["grad"].data_ptr(), out["energy"].data_ptr(), out["forces"].data_ptr(), out["virial"].data_ptr(),
            ctypes.c_void_p(st), ctypes.byref(h)))
        return Sus2Graph(h, [batch, e0, f0, s0], out)

    def tune(self, precision, eb_atoms=2, grid_mult=0):
        """Launch shape of the per-atom EFS kernels (see sus2_tune_set)."""
        _lib.check(_lib.lib().sus2_tune_set(self._h, PRECISION[precision], int(eb_atoms), int(grid_mult)))

    def messages(self, precision, mode=None):
        """Gate-adjoint message mode of the batch EFS ("store", "recompute",
        "split" or "auto", see sus2_tune_messages); sets it when given, returns
        the mode set."""
        names = ["store", "recompute", "split", "auto"]
        cur = ctypes.c_int()
        code = -1 if mode is None else names.index(mode)
        _lib.check(_lib.lib().sus2_tune_messages(self._h, PRECISION[precision], code, ctypes.byref(cur)))
        return names[cur.value]

    def get_tune(self, precision):
        a, b = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().sus2_tune_get(self._h, PRECISION[precision], ctypes.byref(a), ctypes.byref(b)))
        return dict(eb_atoms=a.value, grid_mult=b.value)

    def autotune(self, batch, precision="fp32", reps=3, stream=None):
        """Time every launch shape on this batch and keep the fastest; returns
        (chosen shape, {(eb_atoms, grid_mult): median ms})."""
        import torch
        dev = batch.pos.device
        best = (ctypes.c_int * 2)()
        ms = np.zeros(12)
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_autotune(self._h, PRECISION[precision], batch.nframes, batch.natoms,
                                            batch.off.data_ptr(), batch.types.data_ptr(), batch.pos.data_ptr(),
                                            batch.cell.data_ptr(), batch.periodic.data_ptr(), int(reps),
                                            ctypes.c_void_p(st), best, ms.ctypes.data_as(_lib._dp)))
        table = {((1, 2, 4)[c // 4], (0, 1, 2, 4)[c % 4]): float(ms[c]) for c in range(12)}
        return dict(eb_atoms=best[0], grid_mult=best[1]), table

    def train_bfgs(self, batch, e0, f0, s0=None, weighting="structures", wE=1.0, wF=0.01, wS=0.0,
                   precision="fp32", max_iter=50, gtol=0.0, stream=None):
        """BFGS on the training objective (host FP64 dense inverse Hessian and
        strong-Wolfe line search, GPU loss and gradient); the model's
        coefficients end at the last accepted point.  Returns (loss trace,
        steps taken)."""
        import torch
        dev = batch.pos.device
        mode = {"vibrations": 0, "structures": 1, "molecules": 2}.get(weighting, weighting)
        e0 = e0.to(device=dev, dtype=torch.float64).contiguous()
        f0 = f0.to(device=dev, dtype=torch.float64).contiguous()
        s0 = (torch.zeros((batch.nframes, 6), dtype=torch.float64, device=dev) if s0 is None
              else s0.to(device=dev, dtype=torch.float64).contiguous())
        trace = np.zeros(max_iter + 1)
        iters = ctypes.c_int()
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_train_bfgs(self._h, PRECISION[precision], batch.nframes, batch.natoms,
                                              batch.off.data_ptr(), batch.types.data_ptr(), batch.pos.data_ptr(),
                                              batch.cell.data_ptr(), batch.periodic.data_ptr(), e0.data_ptr(),
                                              f0.data_ptr(), s0.data_ptr(), int(mode), float(wE), float(wF),
                                              float(wS), int(max_iter), float(gtol),
                                              trace.ctypes.data_as(_lib._dp), ctypes.byref(iters),
                                              ctypes.c_void_p(st)))
        return trace[:iters.value + 1], iters.value

    def train_linear(self, batch, e0, f0, s0=None, weighting="structures", wE=1.0, wF=0.01, wS=0.0,
                     precision="fp64", ridge=0.0, stream=None):
        """Linear readout solve (species constants and per-element readout
        weights, the reference's TrainLinear); returns the objective at the
        solution, which replaces those coefficients of the model."""
        import torch
        dev = batch.pos.device
        mode = {"vibrations": 0, "structures": 1, "molecules": 2}.get(weighting, weighting)
        e0 = e0.to(device=dev, dtype=torch.float64).contiguous()
        f0 = f0.to(device=dev, dtype=torch.float64).contiguous()
        s0 = (torch.zeros((batch.nframes, 6), dtype=torch.float64, device=dev) if s0 is None
              else s0.to(device=dev, dtype=torch.float64).contiguous())
        obj = ctypes.c_double()
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_train_linear(self._h, PRECISION[precision], batch.nframes, batch.natoms,
                                                batch.off.data_ptr(), batch.types.data_ptr(), batch.pos.data_ptr(),
                                                batch.cell.data_ptr(), batch.periodic.data_ptr(), e0.data_ptr(),
                                                f0.data_ptr(), s0.data_ptr(), int(mode), float(wE), float(wF),
                                                float(wS), float(ridge), ctypes.byref(obj), ctypes.c_void_p(st)))
        return obj.value

    def train_lbfgs(self, batch, e0, f0, s0=None, weighting="structures", wE=1.0, wF=0.01, wS=0.0,
                    precision="fp32", max_iter=50, history=20, gtol=0.0, stream=None):
        """L-BFGS on the training objective (host FP64 optimiser, GPU loss and
        gradient); the model's coefficients end at the last accepted point.
        Returns (loss trace, steps taken)."""
        import torch
        dev = batch.pos.device
        mode = {"vibrations": 0, "structures": 1, "molecules": 2}.get(weighting, weighting)
        e0 = e0.to(device=dev, dtype=torch.float64).contiguous()
        f0 = f0.to(device=dev, dtype=torch.float64).contiguous()
        s0 = (torch.zeros((batch.nframes, 6), dtype=torch.float64, device=dev) if s0 is None
              else s0.to(device=dev, dtype=torch.float64).contiguous())
        trace = np.zeros(max_iter + 1)
        iters = ctypes.c_int()
        st = _stream_of(stream, dev)
        _lib.check(_lib.lib().sus2_train_lbfgs(self._h, PRECISION[precision], batch.nframes, batch.natoms,
                                               batch.off.data_ptr(), batch.types.data_ptr(), batch.pos.data_ptr(),
                                               batch.cell.data_ptr(), batch.periodic.data_ptr(), e0.data_ptr(),
                                               f0.data_ptr(), s0.data_ptr(), int(mode), float(wE), float(wF),
                                               float(wS), int(max_iter), int(history), float(gtol),
                                               _p(trace), ctypes.byref(iters), ctypes.c_void_p(st)))
        return trace[:iters.value + 1], iters.value


class Sus2Graph:
    """A batch's EFS or objective + gradient captured in a HIP graph
    (sus2_graph_*_create); launch() replays it on a stream.  Inputs and
    outputs are the device tensors given at creation: rewrite them in place
    between launches.  overflow() waits for the stream and reports (and
    clears) whether a launch outgrew the neighbour capacity sized at creation,
    in which case its outputs are invalid and the graph must be re-created."""

    def __init__(self, handle, keep, outputs):
        self._g = handle
        self._keep = keep  # tensors whose device pointers the graph holds
        self.outputs = outputs

    def launch(self, stream=None):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().sus2_graph_launch(self._g, ctypes.c_void_p(st)))
        return self.outputs

    def overflow(self, stream=None):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        v = ctypes.c_int()
        _lib.check(_lib.lib().sus2_graph_overflow(self._g, ctypes.c_void_p(st), ctypes.byref(v)))
        return bool(v.value)

    def close(self):
        if getattr(self, "_g", None) and self._g.value:
            _lib.lib().sus2_graph_free(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _str_weights_from_virial_grad(gV):
    """str_w (row-major 3x3 per frame) that turns sum_e w_e . g_e into
    sum_k gV_k V_k for the PlusStress Voigt components V = -sum d (x) g
    (xx yy zz xy xz yz = -sum d_x g_x, d_y g_y, d_z g_z, d_x g_y, d_x g_z, d_y g_z)."""
    import torch
    nf = gV.shape[0]
    sw = torch.zeros((nf, 3, 3), dtype=torch.float64, device=gV.device)
    # w_e[q] = sum_b sw[q][b] d_b contracts with g_q: V_ab = -sum d_a g_b -> sw[b][a] = -gV_ab
    for k, (a, b) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
        sw[:, b, a] = -gV[:, k]
    return sw


class EFSFunction:
    """torch.autograd.Function over the coefficient vector: forward = batch
    EFS on the GPU, backward = the hand-written parameter adjoint
    (sus2_efs_grad), so that torch optimisers and losses can train the model.

        E, F, V = EFSFunction.apply(theta, model, batch, "fp32")

    theta is an FP64 tensor [ncoeff] (host or device); E [nf], F [N,3],
    V [nf,6] are FP64 device tensors.  Gradients flow to theta only."""

    _fn = None

    @classmethod
    def apply(cls, theta, model, batch, precision="fp32"):
        return cls._function().apply(theta, model, batch, precision)

    @classmethod
    def _function(cls):
        if cls._fn is None:
            import torch

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta, model, batch, precision):
                    model.set_coeffs(theta.detach().to("cpu", torch.float64).numpy())
                    out = model.efs_device(batch, precision)
                    ctx.model, ctx.batch, ctx.precision = model, batch, precision
                    ctx.theta_device = theta.device
                    return out["energy"], out["forces"], out["virial"]

                @staticmethod
                def backward(ctx, gE, gF, gV):
                    b = ctx.batch
                    dev = b.pos.device
                    gE = torch.zeros(b.nframes, dtype=torch.float64, device=dev) if gE is None else gE.double()
                    gF = torch.zeros((b.natoms, 3), dtype=torch.float64, device=dev) if gF is None else gF.double()
                    gV = torch.zeros((b.nframes, 6), dtype=torch.float64, device=dev) if gV is None else gV.double()
                    counts = (b.off[1:] - b.off[:-1]).to(torch.long)
                    ene_w = torch.repeat_interleave(gE, counts)
                    grad = ctx.model.efs_grad_device(b, ene_w, gF, _str_weights_from_virial_grad(gV),
                                                     ctx.precision)
                    return grad.to(ctx.theta_device), None, None, None

            cls._fn = _F
        return cls._fn


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
