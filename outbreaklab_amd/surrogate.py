"""Host side of the microphysics surrogate (no reference counterpart): the network's parameters,
scales and envelope, GPU training from the model's sampled (state, tendency) pairs, and
evaluation. The kernels are in csrc/k_mlp.hip."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

NIN, NOUT, IN, H, OUT = 14, 9, 16, 128, 16
ROW = 24  # floats per sampled row: 14 inputs, 9 tendencies, level
OUT_NAMES = ["dT", "dqv", "dqc", "dqi", "dqr", "dqs", "dqg", "dNr", "dNi"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def init_params(seed=0):
    """He-normal weights, zero biases, in the device layout W1 b1 W2 b2 W3 b3 W4 b4 (W[k][c])."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in ((IN, H), (H, H), (H, H), (H, OUT)):
        parts += [(rng.standard_normal((a, b)) * np.sqrt(2.0 / a)).ravel(), np.zeros(b)]
    return np.concatenate(parts).astype(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, (a, b) in (("1", (IN, H)), ("2", (H, H)), ("3", (H, H)), ("4", (H, OUT))):
        out["W" + name] = p[o:o + a * b].reshape(a, b)
        o += a * b
        out["b" + name] = p[o:o + b]
        o += b
    return out


class Surrogate:
    def __init__(self, seed=0):
        h = C.c_void_p()
        _lib.check(_lib.lib().obl_mlp_create(C.byref(h)))
        self.h = h
        self.n = _lib.lib().obl_mlp_nparams()
        self.set_params(init_params(seed))
        self.out_scale = np.ones(NOUT, np.float32)
        self.in_lo = None
        self.in_hi = None

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _lib.check(_lib.lib().obl_mlp_set_params(self.h, _p(p), p.size), self.h, mlp=True)

    def params(self):
        p = np.empty(self.n, np.float32)
        _lib.check(_lib.lib().obl_mlp_get_params(self.h, _p(p), self.n), self.h, mlp=True)
        return p

    def set_scales(self, out_scale, in_lo=None, in_hi=None):
        self.out_scale = np.ascontiguousarray(out_scale, np.float32)
        lo = None if in_lo is None else np.ascontiguousarray(in_lo, np.float32)
        hi = None if in_hi is None else np.ascontiguousarray(in_hi, np.float32)
        self.in_lo, self.in_hi = lo, hi
        _lib.check(_lib.lib().obl_mlp_set_scales(self.h, _p(self.out_scale), None if lo is None else _p(lo),
                                                 None if hi is None else _p(hi)), self.h, mlp=True)

    def fit_scales(self, rows):
        """Tendency scales = RMS of each output (floored), envelope = min/max of each input."""
        rows = np.asarray(rows, np.float64)
        s = np.sqrt(np.mean(rows[:, NIN:NIN + NOUT] ** 2, axis=0))
        s = np.where(s > 0, s, 1.0)
        self.set_scales(s, rows[:, :NIN].min(0), rows[:, :NIN].max(0))

    def set_dead_zone(self, frac):
        """Absent species are created at inference only above frac * their output scale."""
        _lib.check(_lib.lib().obl_mlp_set_dead_zone(self.h, float(frac)), self.h, mlp=True)

    def set_hand_min(self, qmin):
        """Only points holding at least qmin kg/kg of condensate are handed to the network at inference; thinner
        cloud runs the plain block."""
        _lib.check(_lib.lib().obl_mlp_set_hand_min(self.h, float(qmin)), self.h, mlp=True)

    def load_rows(self, rows):
        """Copy host rows (n x stride float32) to device memory owned by the network."""
        rows = np.ascontiguousarray(rows, np.float32)
        _lib.check(_lib.lib().obl_mlp_load_rows(self.h, _p(rows), rows.shape[0], rows.shape[1]), self.h, mlp=True)

    def train(self, rows_dev_ptr=None, nrows=0, steps=100, batch=4096, lr=1e-3, seed=1, quant=True, stride=ROW):
        """Adam steps on device rows (a device pointer, e.g. the model's sample buffer), or on the
        rows of load_rows when rows_dev_ptr is None. Returns the mean training loss."""
        loss = C.c_float()
        ptr = None if rows_dev_ptr is None else C.c_void_p(rows_dev_ptr)
        _lib.check(_lib.lib().obl_mlp_train(self.h, ptr, nrows, stride, steps, batch, lr, seed,
                                            1 if quant else 0, C.byref(loss)), self.h, mlp=True)
        return loss.value

    def eval(self, rows, quant=True, grads=False):
        rows = np.ascontiguousarray(rows, np.float32)
        n, stride = rows.shape
        loss = C.c_float()
        g = np.zeros(self.n, np.float32) if grads else None
        y = np.zeros((n, OUT), np.float32)
        _lib.check(_lib.lib().obl_mlp_eval(self.h, _p(rows), n, stride, 1 if quant else 0, C.byref(loss),
                                           None if g is None else _p(g), _p(y)), self.h, mlp=True)
        return loss.value, g, y

    def infer(self, rows):
        """Raw outputs (n x 16, before output scaling) of the inference form the step uses."""
        rows = np.ascontiguousarray(rows, np.float32)
        n, stride = rows.shape
        y = np.zeros((n, OUT), np.float32)
        _lib.check(_lib.lib().obl_mlp_infer(self.h, _p(rows), n, stride, _p(y)), self.h, mlp=True)
        return y

    def destroy(self):
        if getattr(self, "h", None) is not None:
            _lib.lib().obl_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def teach(model, dt, steps, capacity=2_000_000, p_active=0.5, p_clear=0.0):
    """Run `steps` plain steps with teacher sampling of the microphysics block on; return the
    sampled rows (n x 24: 14 inputs, 9 tendencies, level) and switch sampling off again."""
    model.sample_micro(capacity, p_active, p_clear)
    for _ in range(steps):
        model.step(dt)
    model.sync()
    rows = model.micro_samples(copy=True)
    model.sample_micro(0)
    return canonical_order(rows)


def canonical_order(rows):
    """The teacher rows sorted by a 64-bit key of their bit patterns. The GPU appends sampled rows with one atomic per
    wave, so their order changes from run to run while the set of rows does not; training draws its batches by row
    index, so without a fixed order two runs of the same seeds train different networks."""
    u = np.ascontiguousarray(rows, dtype=np.float32).view(np.uint32)
    key = np.zeros(u.shape[0], np.uint64)
    mult = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        for c in range(u.shape[1]):
            key = key * mult + u[:, c].astype(np.uint64)
    return np.ascontiguousarray(rows[np.argsort(key, kind="stable")])


NR_IN, NR_OUT = 9, NIN + 7  # Nr input and dNr output columns of a teacher row


def targets(rows):
    """Teacher rows with the network's targets: the number-concentration output becomes the log
    ratio log((Nr + dNr + 1) / (Nr + 1)). Nr changes mostly multiplicatively (self-collection and
    breakup, evaporation), over orders of magnitude; a linear dNr scaled to its largest values
    leaves errors larger than small Nr themselves."""
    t = np.array(rows, dtype=np.float32, copy=True)
    n0 = np.maximum(t[:, NR_IN].astype(np.float64), 0.0)
    n1 = np.maximum(n0 + t[:, NR_OUT], 0.0)
    t[:, NR_OUT] = np.log((n1 + 1.0) / (n0 + 1.0))
    return t


def fit(net, rows, steps, batch=4096, lr=2e-3, seed=1, chunk=500, quant=True):
    """Targets (see `targets`), scales and envelope from the teacher rows, then `steps` Adam steps on
    the GPU with the learning rate halved for each of the last three quarters. Returns the mean
    training loss of each chunk."""
    rows = targets(rows)
    net.fit_scales(rows)
    net.load_rows(rows)
    losses = []
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        frac = done / max(steps, 1)
        rate = lr * 0.5 ** int(4 * frac)
        losses.append(net.train(steps=k, batch=batch, lr=rate, seed=seed + done, quant=quant))
        done += k
    return losses

