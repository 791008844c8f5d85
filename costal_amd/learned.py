"""Learned admission thresholds for the fused kNN scan.

A 3-layer MLP (d -> 64 -> 64 -> 1, ReLU) on L2-normalized rows predicts u = log(1 - s_k + 1e-6)
for a row's k-th neighbor cosine s_k.  Inside a kNN call the library turns it into a threshold
t = 1 - exp(u + offset) - 2 eps; candidates scoring at or below t skip the candidate-list insert.
``offset`` is calibrated on held-out rows so that the threshold lies below s_k for a chosen fraction
of rows; every other row is caught by the scan's certificate and recomputed, so the result is exact
whatever the model predicts.  ``offset2`` (the largest held-out residual plus a margin) gives the
rows left with fewer than k candidates a second, lower threshold for their rescan, so that it
prunes like the first pass; rows short again are recomputed by the exact row scan.

Training data come from the exact method itself: the exact kNN of a training sample gives the
labels.  When the sample is smaller than the data the model will serve, the k-th neighbor of the
larger set sits about where the (k * n_train / n_target)-th neighbor of the sample does (neighbor
counts scale with density), so that neighbor rank is the label.

The network is trained on the GPU by the library's own kernels (costal_mlp_grad_device,
costal_mlp_adam_device: deterministic fixed-order reductions) and evaluated in BF16 inside the kNN
call.  Device buffers are torch tensors (plumbing only).
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _costal, _lib

H = 64
EPS_LABEL = 1e-6


def param_count(d: int) -> int:
    return H * d + H + H * H + H + H + 1


def init_params(d: int, seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases (float32, the library's flat layout)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, np.sqrt(2.0 / d), size=(H, d))
    w2 = rng.normal(0.0, np.sqrt(2.0 / H), size=(H, H))
    w3 = rng.normal(0.0, np.sqrt(1.0 / H), size=H)
    return np.concatenate([w1.ravel(), np.zeros(H), w2.ravel(), np.zeros(H), w3, [0.0]]).astype(np.float32)


def label_transform(s: np.ndarray) -> np.ndarray:
    return np.log(np.maximum(1.0 - np.asarray(s, dtype=np.float64), 0.0) + EPS_LABEL).astype(np.float32)


@dataclass
class ThresholdModel:
    d: int
    k: int
    params: np.ndarray
    offset: float
    meta: dict = field(default_factory=dict)
    offset2: float | None = None  # second-chance offset (None: rescan short rows without threshold)

    def save(self, path: str) -> None:
        np.savez(path, d=self.d, k=self.k, params=self.params, offset=self.offset,
                 offset2=self.offset if self.offset2 is None else self.offset2)

    @classmethod
    def load(cls, path: str) -> "ThresholdModel":
        f = np.load(path, allow_pickle=False)
        o2 = float(f["offset2"]) if "offset2" in f.files else None
        return cls(int(f["d"]), int(f["k"]), f["params"].astype(np.float32), float(f["offset"]), offset2=o2)


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("training the threshold model needs a GPU")
    return torch


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def predict_raw(params: np.ndarray, x: np.ndarray, *, bf16: bool = True) -> np.ndarray:
    """The model output u for every row of x (GPU)."""
    torch = _torch()
    L = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    dx = torch.from_numpy(x).cuda()
    dp = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).cuda()
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _check(L.costal_mlp_predict_device(_ptr(dx), n, d, _ptr(dp), int(bf16), 0.0, 0.0, 1, _ptr(out),
                                       C.c_void_p(stream)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _check(rc: int) -> None:
    if rc != _lib.OK:
        raise _lib.CostalError(rc, _lib.load().costal_mlp_last_error().decode())


def gpu_gradient(params: np.ndarray, x: np.ndarray, target: np.ndarray):
    """(loss, gradient) of the mean squared error over one batch, computed by the GPU kernels."""
    torch = _torch()
    L = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    dx = torch.from_numpy(x).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(target, dtype=np.float32)).cuda()
    dp = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).cuda()
    act = torch.empty(int(L.costal_mlp_act_floats(n)), dtype=torch.float32, device="cuda")
    grad = torch.empty(param_count(d), dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _check(L.costal_mlp_grad_device(_ptr(dx), _ptr(dt), n, d, _ptr(dp), _ptr(act), _ptr(grad), _ptr(loss),
                                    C.c_void_p(stream)))
    torch.cuda.synchronize()
    return float(loss.cpu()[0]), grad.cpu().numpy()


def fit(x: np.ndarray, u: np.ndarray, *, epochs: int = 60, batch: int = 512, lr: float = 3e-3,
        seed: int = 0, params: np.ndarray | None = None, log=None):
    """Adam with a cosine learning-rate decay on the mean squared error of u, all on the GPU;
    returns (params, per-epoch losses)."""
    torch = _torch()
    L = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    p0 = init_params(d, seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
    P = param_count(d)
    dx = torch.from_numpy(x).cuda()
    du = torch.from_numpy(np.ascontiguousarray(u, dtype=np.float32)).cuda()
    dp = torch.from_numpy(p0.copy()).cuda()
    m = torch.zeros(P, dtype=torch.float32, device="cuda")
    v = torch.zeros(P, dtype=torch.float32, device="cuda")
    grad = torch.empty(P, dtype=torch.float32, device="cuda")
    act = torch.empty(int(L.costal_mlp_act_floats(batch)), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    bx = torch.empty((batch, d), dtype=torch.float32, device="cuda")
    bu = torch.empty(batch, dtype=torch.float32, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    losses = []
    step = 0
    total = max(1, epochs * (n // batch))
    for ep in range(epochs):
        order = torch.randperm(n, device="cuda", generator=gen)
        ep_loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        nb = 0
        for b0 in range(0, n - batch + 1, batch):
            sel = order[b0:b0 + batch]
            torch.index_select(dx, 0, sel, out=bx)
            torch.index_select(du, 0, sel, out=bu)
            _check(L.costal_mlp_grad_device(_ptr(bx), _ptr(bu), batch, d, _ptr(dp), _ptr(act), _ptr(grad),
                                            _ptr(loss), stream))
            lr_t = 0.5 * lr * (1.0 + np.cos(np.pi * step / total))
            step += 1
            _check(L.costal_mlp_adam_device(P, _ptr(dp), _ptr(grad), _ptr(m), _ptr(v), lr_t, 0.9, 0.999, 1e-8,
                                            step, stream))
            ep_loss += loss
            nb += 1
        losses.append(float(ep_loss.cpu()[0]) / max(nb, 1))
        if log:
            log(f"epoch {ep + 1}/{epochs} mse {losses[-1]:.5f}")
    torch.cuda.synchronize()
    return dp.cpu().numpy(), losses


def train(x_train: np.ndarray, k: int, *, n_target: int | None = None, epochs: int = 60,
          batch: int = 512, lr: float = 3e-3, holdout: float = 0.1, quantile: float = 0.995,
          seed: int = 0, log=None) -> ThresholdModel:
    """Train a threshold model on a sample of the data it will serve.

    Labels are the exact k-th neighbor similarities of the sample (GPU exact kNN); with
    ``n_target`` the neighbor rank is scaled to k * n_train / n_target.  ``quantile`` of the held-out
    rows get a threshold below their k-th similarity (the others fall back to the plain scan)."""
    x_train = np.ascontiguousarray(x_train, dtype=np.float32)
    n, d = x_train.shape
    k_eff = k if not n_target else max(1, int(round(k * n / n_target)))
    t0 = time.perf_counter()
    (idx, cos, _), _ = _costal._knn_exact(x_train, k_eff)
    s = cos[:, k_eff - 1]
    u = label_transform(s)
    t_label = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    n_hold = max(1, int(holdout * n))
    hold, tr = perm[:n_hold], perm[n_hold:]
    t0 = time.perf_counter()
    params, losses = fit(x_train[tr], u[tr], epochs=epochs, batch=min(batch, len(tr)), lr=lr, seed=seed,
                         log=log)
    t_fit = time.perf_counter() - t0
    pred = predict_raw(params, x_train[hold], bf16=True)
    resid = u[hold].astype(np.float64) - pred.astype(np.float64)
    offset = float(np.quantile(resid, quantile))
    # second chance: above every held-out residual by a margin of 0.1 in log(1 - s) (10% in 1 - s)
    offset2 = max(offset, float(resid.max())) + 0.1
    meta = {
        "n_train": n,
        "k_label": k_eff,
        "epochs": epochs,
        "final_mse": losses[-1] if losses else None,
        "holdout_rmse": float(np.sqrt(np.mean(resid ** 2))),
        "quantile": quantile,
        "label_seconds": t_label,
        "fit_seconds": t_fit,
        "losses": losses,
    }
    return ThresholdModel(d=d, k=k, params=params, offset=offset, meta=meta, offset2=offset2)
