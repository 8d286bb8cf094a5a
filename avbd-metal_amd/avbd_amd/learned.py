"""WarmStartMLP: the learned warm start's network over the C ABI (avbd.h).

A per-body MLP 24 -> 128 -> 128 -> 3 (particles) or 6 (rigid bodies), ReLU,
BF16 weights and activations, FP32 accumulation, whose forward, backward and
Adam steps run as hand-written gfx950 kernels.  It predicts a step's
converged correction relative to the adaptive warm-start guess in units of
|g| dt^2; attach it with GPUSolver.setWarmStart.  The reference has no
learned warm start.
"""
import ctypes as C

import numpy as np

from . import _capi

NUM_PARAMS = _capi.MLP_PARAMS
INPUTS = _capi.MLP_INPUTS
OUTPUTS = _capi.MLP_OUTPUTS


def layout(outputs=OUTPUTS):
    """Shapes of the flat parameter array: W1[24][128], b1, W2[128][128], b2,
    W3[128][outputs], b3 (W[in][out])."""
    return (("W1", (INPUTS, 128)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("W3", (128, outputs)),
            ("b3", (outputs,)))


LAYOUT = layout()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _rows(a, width):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != width:
        raise ValueError(f"expected an [n, {width}] array, got {a.shape}")
    return a


def split_params(p, outputs=OUTPUTS):
    """Flat parameter array -> dict of named arrays (views)."""
    out, o = {}, 0
    for name, shape in layout(outputs):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


class WarmStartMLP:
    def __init__(self, seed=0, outputs=OUTPUTS):
        """outputs 3: a particle network; 6: a rigid-body network."""
        self._lib = _capi.load()
        h = C.c_void_p()
        st = self._lib.avbd_mlp_create(C.c_uint64(seed), int(outputs), C.byref(h))
        if st != _capi.OK:
            raise _capi.AVBDError(st, "avbd_mlp_create")
        self._h = h
        o, n = C.c_int32(), C.c_int32()
        self._lib.avbd_mlp_info(h, C.byref(o), C.byref(n))
        self.outputs, self.numParams = o.value, n.value
        _capi.track("network", self)

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.avbd_mlp_destroy(self._h)
            self._h = None

    def _check(self, st, what):
        if st != _capi.OK:
            raise _capi.AVBDError(st, what)

    @property
    def params(self):
        p = np.zeros(self.numParams, np.float32)
        self._check(self._lib.avbd_mlp_get_params(self._h, _fp(p)), "get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (self.numParams,):
            raise ValueError(f"expected {self.numParams} parameters")
        self._check(self._lib.avbd_mlp_set_params(self._h, _fp(p)), "set_params")

    def forward(self, x):
        x = _rows(x, INPUTS)
        y = np.zeros((len(x), self.outputs), np.float32)
        self._check(self._lib.avbd_mlp_forward(self._h, _fp(x), len(x), _fp(y)), "forward")
        return y

    def lossGrad(self, x, t):
        """(loss, grad) of sum |y - t|^2 / (outputs n) over one batch of rows."""
        x, t = _rows(x, INPUTS), _rows(t, self.outputs)
        loss = C.c_float()
        g = np.zeros(self.numParams, np.float32)
        self._check(self._lib.avbd_mlp_loss_grad(self._h, _fp(x), _fp(t), len(x), C.byref(loss), _fp(g)),
                    "loss_grad")
        return loss.value, g

    def train(self, x, t, epochs=10, batch=4096, lr=1e-3, seed=0):
        """Adam over shuffled batches on the GPU; returns the mean loss per epoch."""
        x, t = _rows(x, INPUTS), _rows(t, self.outputs)
        if len(x) != len(t):
            raise ValueError("features and targets differ in length")
        losses = np.zeros(int(epochs), np.float32)
        self._check(self._lib.avbd_mlp_train(self._h, _fp(x), _fp(t), len(x), int(epochs), int(batch), float(lr),
                                             C.c_uint64(seed), _fp(losses)), "train")
        return losses

    def timeForward(self, n, reps=20):
        """Average microseconds of one forward launch over n rows."""
        us = C.c_float()
        self._check(self._lib.avbd_mlp_time_forward(self._h, int(n), int(reps), C.byref(us)), "time_forward")
        return us.value


def collect_samples(scenes, frames, tol=2.0e-4, max_iterations=200, check_every=10, rigid=False, **solver_kw):
    """Training pairs from the plain solver: each scene is stepped `frames`
    times, every frame iterated until residual <= tol (at most
    max_iterations sweeps); returns (features [n,24], corrections [n,3]) of
    every dynamic particle (rigid=True: [n,6] of every dynamic rigid body)
    and frame, and the mean sweeps per frame."""
    from .solver import GPUSolver

    xs, ts, its = [], [], []
    for scene in scenes:
        g = GPUSolver(scene, **solver_kw)
        g.setWarmStart(None, log=True)
        for _ in range(frames):
            n, _ = g.stepToResidual(tol, max_iterations, check_every)
            its.append(n)
            x, t, _ = g.warmStartSamples(rigid=rigid)
            xs.append(x)
            ts.append(t)
        g.close()
    return np.concatenate(xs), np.concatenate(ts), float(np.mean(its))
