"""ctypes binding of liblymph_mlp.so (include/lymph_mlp.h): a fully connected
GELU network trained and evaluated on the GPU with BF16 MFMA kernels and
FP32 master weights.  Fails loudly when the library or the GPU is missing."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

EXPORTS = [
    "lymph_mlp_last_error", "lymph_mlp_create", "lymph_mlp_destroy", "lymph_mlp_n_params",
    "lymph_mlp_set_params", "lymph_mlp_get_params", "lymph_mlp_set_data", "lymph_mlp_train",
    "lymph_mlp_gradients", "lymph_mlp_predict", "lymph_mlp_time_predict",
]


class MlpError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblymph_mlp.so")
        if not os.path.exists(path):
            raise MlpError(-1, "load", f"{path} is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        vp, i, f, d = C.c_void_p, C.c_int, C.c_float, C.c_double
        sig = {
            "lymph_mlp_last_error": (C.c_char_p, []),
            "lymph_mlp_create": (i, [i, vp, i, vp]),
            "lymph_mlp_destroy": (i, [vp]),
            "lymph_mlp_n_params": (i, [vp, vp]),
            "lymph_mlp_set_params": (i, [vp, vp]),
            "lymph_mlp_get_params": (i, [vp, vp]),
            "lymph_mlp_set_data": (i, [vp, i, vp, vp]),
            "lymph_mlp_train": (i, [vp, i, i, vp, f, f, f, f, vp]),
            "lymph_mlp_gradients": (i, [vp, i, vp, vp, vp, vp]),
            "lymph_mlp_predict": (i, [vp, i, vp, vp]),
            "lymph_mlp_time_predict": (i, [vp, i, i, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _check(rc, where):
    if rc != 0:
        raise MlpError(rc, where, lib().lymph_mlp_last_error().decode())


def init_params(widths, seed):
    """torch.nn.Linear's default initialisation: W, b ~ U(-1/sqrt(n_in), 1/sqrt(n_in))."""
    rng = np.random.default_rng(seed)
    out = []
    for k, n in zip(widths[:-1], widths[1:]):
        s = 1.0 / np.sqrt(k)
        out.append(rng.uniform(-s, s, n * k).astype(np.float32))
        out.append(rng.uniform(-s, s, n).astype(np.float32))
    return np.concatenate(out)


class Mlp:
    """Network of len(widths)-1 linear layers with GELU between them."""

    def __init__(self, widths, max_batch=256, seed=0):
        self.widths = [int(w) for w in widths]
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        w = np.ascontiguousarray(self.widths, np.int32)
        _check(lib().lymph_mlp_create(len(self.widths) - 1, _p(w), self.max_batch, C.byref(h)),
               "lymph_mlp_create")
        self.h = h
        n = np.zeros(1, np.int64)
        _check(lib().lymph_mlp_n_params(h, _p(n)), "lymph_mlp_n_params")
        self.n_params = int(n[0])
        self.set_params(init_params(self.widths, seed))

    def close(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.lymph_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"{p.size} parameters, expected {self.n_params}")
        _check(lib().lymph_mlp_set_params(self.h, _p(p)), "lymph_mlp_set_params")

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(lib().lymph_mlp_get_params(self.h, _p(p)), "lymph_mlp_get_params")
        return p

    def set_data(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        assert x.shape[1] == self.widths[0] and y.shape[1] == self.widths[-1] and x.shape[0] == y.shape[0]
        _check(lib().lymph_mlp_set_data(self.h, x.shape[0], _p(x), _p(y)), "lymph_mlp_set_data")
        self.n_data = x.shape[0]

    def train(self, order, batch, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        order = np.ascontiguousarray(order, np.int32).ravel()
        steps = order.size // batch
        losses = np.zeros(steps, np.float32)
        _check(lib().lymph_mlp_train(self.h, steps, batch, _p(order), lr, beta1, beta2, eps,
                                     _p(losses)), "lymph_mlp_train")
        return losses

    def gradients(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = np.zeros(1, np.float32)
        _check(lib().lymph_mlp_gradients(self.h, x.shape[0], _p(x), _p(y), _p(g), _p(loss)),
               "lymph_mlp_gradients")
        return float(loss[0]), g

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.widths[-1]), np.float32)
        _check(lib().lymph_mlp_predict(self.h, x.shape[0], _p(x), _p(y)), "lymph_mlp_predict")
        return y

    def time_predict(self, n, reps=10):
        ms = C.c_double()
        _check(lib().lymph_mlp_time_predict(self.h, n, reps, C.byref(ms)), "lymph_mlp_time_predict")
        return ms.value


if __name__ == "__main__":
    # Training-step timing of the surrogate's network (widths 8-512-512-512-
    # 38400, batch 256) on random data: python -m lymph3d_apsc_amd.mlp
    import time
    widths, batch = [8, 512, 512, 512, 38400], 256
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1024, widths[0])).astype(np.float32)
    y = rng.standard_normal((1024, widths[-1])).astype(np.float32)
    with Mlp(widths, max_batch=batch, seed=0) as net:
        net.set_data(x, y)
        order = np.concatenate([rng.permutation(1024) for _ in range(10)]).astype(np.int32)[:batch * 40]
        net.train(order[:batch * 4], batch, 1e-3)
        t = time.perf_counter()
        net.train(order, batch, 1e-3)
        dt = time.perf_counter() - t
        flops = 2.0 * batch * sum(k * n for k, n in zip(widths[:-1], widths[1:])) * 3 - 2.0 * batch * widths[0] * widths[1]
        print(f"training step {1e3 * dt / 40:.3f} ms ({flops * 40 / dt / 1e12:.1f} TFLOP/s); "
              f"forward of 1024 samples {net.time_predict(1024, 10):.3f} ms", flush=True)
