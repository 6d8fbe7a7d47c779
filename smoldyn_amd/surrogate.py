"""Learned surrogate for the reaction-parameter inversion.

For a reversible reaction whose unbinding radius is a fixed multiple of the
binding radius (the reference's "ratio" product placement), the setup turns
the requested rate into a binding radius by bisection on the reference's
numerical rate function (rxnparam.c:558-583, bindingradius).  The surrogate
replaces the bisection with a 3-64-64-2 MLP evaluated in BF16 on the GPU,
one secant step on the rate function (two where the first leaves the error
above the tolerance), and a check of the result against the
rate function (the judge): a query whose relative rate error exceeds the
tolerance (1e-3) falls back to the bisection.

Network features (see smolamdRxnparamSamples): x = (log(rate*dt/step^3),
log(step), ratio) -> y = (log(a/step), log(b/step)); a is the binding radius,
b = ratio * a the unbinding radius, step = sqrt(2 D dt) the rms step.  The
network is trained on the GPU by the library's own forward, backward and Adam
kernels (surrogate.hip) on samples of the reference's rate function.
"""
import ctypes as C
import time
from pathlib import Path

import numpy as np

from . import lib

RANGES = (0.05, 5.0, 1.0, 3.0, 0.01, 10.0)  # step/a, b/a, step
DATA = Path(__file__).resolve().parent / "data" / "rate_surrogate.npz"
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        P, D, I, L64 = C.c_void_p, C.c_double, C.c_int, C.c_long
        for name, res, args in [
            ("smolamdSurrogateParamCount", I, []),
            ("smolamdSurrogateNew", P, [_fp]),
            ("smolamdSurrogateFree", None, [P]),
            ("smolamdSurrogateGetParams", I, [P, _fp]),
            ("smolamdSurrogateGradient", I, [P, _fp, _fp, I, _fp, _dp]),
            ("smolamdSurrogateTrain", I, [P, _fp, _fp, L64, C.POINTER(C.c_int), I, I, C.c_float, C.c_float, _dp]),
            ("smolamdSurrogatePredict", I, [P, _fp, I, _fp, I, _dp]),
            ("smolamdSetRateSurrogate", I, [_fp, _dp, D]),
            ("smolamdRateSurrogateStats", I, [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
            ("smolamdRxnparamSamples", L64, [L64, C.c_ulonglong, _dp, _dp, _dp]),
            ("smolamdRxnparamInvert", I, [L64, _dp, _dp, _dp, _dp]),
            ("smolamdRxnparamFeatures", I, [L64, _dp, _dp, _dp, _dp]),
            ("smolamdRxnparamPolish", I, [L64, _dp, _dp, _dp, _dp, D, _dp]),
            ("smolamdRxnparamRateError", I, [L64, _dp, _dp, _dp, _dp, _dp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return L


def _d(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(_dp)


def nparam():
    return _lib().smolamdSurrogateParamCount()


def samples(n, seed=1, ranges=RANGES):
    """n samples of the reference's rate function: features X (n,3), targets
    Y (n,2), and the number of redrawn points."""
    X = np.zeros((n, 3))
    Y = np.zeros((n, 2))
    rd = _lib().smolamdRxnparamSamples(n, seed, _d(np.array(ranges, np.float64)), _d(X), _d(Y))
    return X, Y, rd


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (layout W1 b1 W2 b2 W3 b3)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((3, 64), (64, 64), (64, 2)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def rate_error(ratedt, step, ratio, a):
    """Relative error of the reference's rate function at binding radii a."""
    out = np.zeros(len(a))
    _lib().smolamdRxnparamRateError(len(a), _d(ratedt), _d(step), _d(ratio), _d(a), _d(out))
    return out


def invert_numerical(ratedt, step, ratio):
    a = np.zeros(len(ratedt))
    _lib().smolamdRxnparamInvert(len(a), _d(ratedt), _d(step), _d(ratio), _d(a))
    return a


class Net:
    """Device-side network (weights on the GPU)."""

    def __init__(self, params):
        self.L = _lib()
        p = np.ascontiguousarray(params, np.float32)
        self.h = self.L.smolamdSurrogateNew(p.ctypes.data_as(_fp))
        if not self.h:
            raise RuntimeError("surrogate allocation failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.smolamdSurrogateFree(self.h)
            self.h = None

    def params(self):
        p = np.zeros(nparam(), np.float32)
        self.L.smolamdSurrogateGetParams(self.h, p.ctypes.data_as(_fp))
        return p

    def gradient(self, Xs, Ts):
        Xs = np.ascontiguousarray(Xs, np.float32)
        Ts = np.ascontiguousarray(Ts, np.float32)
        g = np.zeros(nparam(), np.float32)
        loss = C.c_double(0)
        rc = self.L.smolamdSurrogateGradient(self.h, Xs.ctypes.data_as(_fp), Ts.ctypes.data_as(_fp), len(Xs),
                                             g.ctypes.data_as(_fp), C.byref(loss))
        if rc:
            raise RuntimeError("surrogate gradient failed")
        return g, loss.value

    def train(self, Xs, Ts, epochs, batch, lr, lr_final, seed=0):
        Xs = np.ascontiguousarray(Xs, np.float32)
        Ts = np.ascontiguousarray(Ts, np.float32)
        n = len(Xs)
        rng = np.random.default_rng(seed)
        perm = np.concatenate([rng.permutation(n) for _ in range(epochs)]).astype(np.int32)
        losses = np.zeros(epochs)
        rc = self.L.smolamdSurrogateTrain(self.h, Xs.ctypes.data_as(_fp), Ts.ctypes.data_as(_fp), n,
                                          perm.ctypes.data_as(C.POINTER(C.c_int)), epochs, batch, lr, lr_final,
                                          _d(losses))
        if rc:
            raise RuntimeError("surrogate training failed")
        return losses

    def predict(self, Xs, bf16=True):
        Xs = np.ascontiguousarray(Xs, np.float32)
        Y = np.zeros((len(Xs), 2), np.float32)
        ms = C.c_double(0)
        rc = self.L.smolamdSurrogatePredict(self.h, Xs.ctypes.data_as(_fp), len(Xs), Y.ctypes.data_as(_fp),
                                            1 if bf16 else 0, C.byref(ms))
        if rc:
            raise RuntimeError("surrogate inference failed")
        return Y, ms.value


class RateSurrogate:
    """Trained network with its feature standardization."""

    def __init__(self, params, norm):
        self.params = np.ascontiguousarray(params, np.float32)
        self.norm = np.ascontiguousarray(norm, np.float64)  # in mean[3], in sd[3], out mean[2], out sd[2]
        self._net = None

    @classmethod
    def load(cls, path=DATA):
        d = np.load(path, allow_pickle=False)
        return cls(d["params"], d["norm"])

    def save(self, path=DATA, **info):
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, params=self.params, norm=self.norm, **info)

    @classmethod
    def train(cls, n=1_000_000, epochs=200, batch=4096, lr=3e-3, lr_final=1e-5, seed=1, progress=None):
        """Train on n samples of the reference's rate function on the GPU.
        The learning rate decays geometrically from lr to lr_final over the
        run; progress(epoch, loss) is called after each epoch."""
        X, Y, _ = samples(n, seed)
        norm = np.concatenate([X.mean(0), X.std(0), Y.mean(0), Y.std(0)])
        Xs = (X - norm[:3]) / norm[3:6]
        Ts = (Y - norm[6:8]) / norm[8:10]
        net = Net(init_params(seed))
        losses = np.zeros(epochs)
        ratio = lr_final / lr
        for e in range(epochs):
            lo, hi = lr * ratio ** (e / epochs), lr * ratio ** ((e + 1) / epochs)
            losses[e] = net.train(Xs, Ts, 1, batch, lo, hi, seed + e)[0]
            if progress:
                progress(e, losses[e])
        m = cls(net.params(), norm)
        m._net = net
        return m, losses

    def net(self):
        if self._net is None:
            self._net = Net(self.params)
        return self._net

    def standardize(self, X):
        return (X - self.norm[:3]) / self.norm[3:6]

    def radius(self, ratedt, step, ratio, bf16=True):
        """Network binding radius (no correction) and the inference time."""
        X = np.zeros((len(ratedt), 3))
        _lib().smolamdRxnparamFeatures(len(ratedt), _d(ratedt), _d(step), _d(ratio), _d(X))
        Y, ms = self.net().predict(self.standardize(X), bf16=bf16)
        return np.asarray(step) * np.exp(Y[:, 0] * self.norm[8] + self.norm[6]), ms

    def invert(self, ratedt, step, ratio, tol=1e-3, bf16=True):
        """Binding radii for the queries: network (GPU), secant steps on the
        rate function, judge; queries the judge rejects go to the
        bisection.  Returns radii, accepted mask and timings (s)."""
        ratedt = np.ascontiguousarray(ratedt, np.float64)
        step = np.ascontiguousarray(step, np.float64)
        ratio = np.ascontiguousarray(ratio, np.float64)
        t0 = time.perf_counter()
        a0, ms = self.radius(ratedt, step, ratio, bf16)
        t1 = time.perf_counter()
        a = np.zeros(len(a0))
        _lib().smolamdRxnparamPolish(len(a), _d(ratedt), _d(step), _d(ratio), _d(a0), tol, _d(a))
        err = rate_error(ratedt, step, ratio, a)
        ok = err <= tol
        t2 = time.perf_counter()
        if not ok.all():
            bad = ~ok
            a[bad] = invert_numerical(ratedt[bad], step[bad], ratio[bad])
        t3 = time.perf_counter()
        return a, ok, dict(network_s=t1 - t0, network_kernel_ms=ms, polish_and_judge_s=t2 - t1, fallback_s=t3 - t2,
                           total_s=t3 - t0)

    def install(self, tol=1e-3):
        """Use the surrogate in smolUpdateSim for ratio-placement reactions."""
        _lib().smolamdSetRateSurrogate(self.params.ctypes.data_as(_fp), _d(self.norm), tol)

    @staticmethod
    def uninstall():
        _lib().smolamdSetRateSurrogate(None, None, 0.0)

    @staticmethod
    def stats():
        a, r = C.c_longlong(0), C.c_longlong(0)
        _lib().smolamdRateSurrogateStats(C.byref(a), C.byref(r))
        return a.value, r.value


def train_main():
    """Train the network on the GPU and store it in the package data."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--out", default=str(DATA))
    args = ap.parse_args()
    t0 = time.perf_counter()
    m, losses = RateSurrogate.train(args.samples, args.epochs, args.batch,
                                    progress=lambda e, l: print(f"epoch {e} loss {l:.3e} "
                                                                f"{time.perf_counter() - t0:.1f}s", flush=True))
    t_train = time.perf_counter() - t0
    Xt, Yt, _ = samples(20000, seed=99)
    ratedt = np.exp(Xt[:, 0] + 3 * Xt[:, 1])
    step = np.exp(Xt[:, 1])
    ratio = Xt[:, 2]
    a_net, _ = m.radius(ratedt, step, ratio, bf16=True)
    err_raw = rate_error(ratedt, step, ratio, a_net)
    a, ok, _ = m.invert(ratedt, step, ratio)
    info = dict(samples=args.samples, epochs=args.epochs, batch=args.batch, train_s=t_train,
                final_loss=float(losses[-1]), raw_bf16_median_rate_error=float(np.median(err_raw)),
                accepted_fraction=float(ok.mean()))
    m.save(args.out, losses=losses)
    print(json.dumps(info))


if __name__ == "__main__":
    train_main()
