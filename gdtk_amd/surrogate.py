"""Learned chemistry surrogate: ctypes host of the network and surrogate entry points
(include/gdtk_hip.h) and the training pipeline.

The surrogate has no counterpart in the reference.  It is trained only on (state, dt,
result) samples produced by this library's own alpha-QSS integrator (gdtk_hip_chem_step_device
over cells of the chemistry workload), with the library's FP32 training kernels, and is
evaluated in BF16 on the matrix cores.  Its step (gdtk_hip_surrogate_step_device) is a
drop-in for the integrator's: cells outside the trained range or where the two networks
disagree are handed to the integrator, and every answer conserves each element's mass.
PyTorch is used here only to hold device arrays.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from .chem import MAX_SPECIES, _check, element_matrix

MAX_FEAT = 11
MAX_ELEM = 4

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_bound = False


class MlpDesc(C.Structure):
    _fields_ = [("n_in", C.c_int), ("hidden", C.c_int), ("n_hidden_layers", C.c_int), ("n_out", C.c_int)]


class SurrogateCfg(C.Structure):
    _fields_ = [("x_mean", C.c_double * MAX_FEAT), ("x_inv_std", C.c_double * MAX_FEAT),
                ("y_scale", C.c_double * MAX_SPECIES), ("x_abs_max", C.c_double), ("disagree_tol", C.c_double),
                ("elem", (C.c_double * MAX_SPECIES) * MAX_ELEM), ("n_elem", C.c_int),
                ("route_ratio", C.c_double), ("route_T", C.c_double)]


def _lib():
    global _bound
    from . import hip
    L = hip.lib()
    if not _bound:
        L.gdtk_hip_mlp_create.argtypes = [C.POINTER(MlpDesc), C.c_ulonglong, C.POINTER(C.c_int)]
        L.gdtk_hip_mlp_destroy.argtypes = [C.c_int]
        L.gdtk_hip_mlp_nparams.argtypes = [C.c_int]
        L.gdtk_hip_mlp_get_params.argtypes = [C.c_int, _fp, C.c_int]
        L.gdtk_hip_mlp_set_params.argtypes = [C.c_int, _fp, C.c_int]
        L.gdtk_hip_mlp_forward.argtypes = [C.c_int, C.c_int, _fp, _fp]
        L.gdtk_hip_mlp_set_bf16_training.argtypes = [C.c_int, C.c_int]
        L.gdtk_hip_mlp_loss_grad.argtypes = [C.c_int, C.c_int, _fp, _fp, _fp, _fp]
        L.gdtk_hip_mlp_train.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float,
                                         C.c_ulonglong, _fp, C.c_int]
        L.gdtk_hip_mlp_infer.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.gdtk_hip_mfma_probe.argtypes = [_fp, _fp, _fp]
        L.gdtk_hip_surrogate_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(SurrogateCfg),
                                                C.POINTER(C.c_int)]
        L.gdtk_hip_surrogate_destroy.argtypes = [C.c_int]
        L.gdtk_hip_surrogate_step_device.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_double] + \
            [C.c_void_p] * 8
        L.gdtk_hip_surrogate_stats.argtypes = [C.c_int, _dp, C.c_int]
        L.gdtk_hip_surrogate_project.argtypes = [C.c_int, C.c_int, _dp, C.c_int] + [C.c_void_p] * 5
        _bound = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mfma_probe(A, B):
    """D = A B through the matrix-core lane layout of the inference kernel (16x32 by 32x16)."""
    A, B = _f32(A), _f32(B)
    D = np.zeros((16, 16), dtype=np.float32)
    _check(_lib().gdtk_hip_mfma_probe(A.ctypes.data_as(_fp), B.ctypes.data_as(_fp), D.ctypes.data_as(_fp)))
    return D


class Mlp:
    """The surrogate network (gdtk_hip_mlp_*): n_in -> hidden x n_hidden (GELU) -> n_out."""

    def __init__(self, n_in, n_out, hidden=256, n_hidden=3, seed=0):
        self.n_in, self.n_out, self.hidden, self.n_hidden = n_in, n_out, hidden, n_hidden
        h = C.c_int(-1)
        _check(_lib().gdtk_hip_mlp_create(C.byref(MlpDesc(n_in, hidden, n_hidden, n_out)), seed, C.byref(h)))
        self.h = h.value
        self.nparams = _check(_lib().gdtk_hip_mlp_nparams(self.h))

    def close(self):
        if getattr(self, "h", -1) >= 0:
            _lib().gdtk_hip_mlp_destroy(self.h)
            self.h = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(_lib().gdtk_hip_mlp_get_params(self.h, p.ctypes.data_as(_fp), self.nparams))
        return p

    def set_params(self, p):
        p = _f32(p)
        _check(_lib().gdtk_hip_mlp_set_params(self.h, p.ctypes.data_as(_fp), self.nparams))

    def forward(self, X):
        X = _f32(X)
        Y = np.zeros((X.shape[0], self.n_out), dtype=np.float32)
        _check(_lib().gdtk_hip_mlp_forward(self.h, X.shape[0], X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp)))
        return Y

    def loss_grad(self, X, Y):
        X, Y = _f32(X), _f32(Y)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float(0.0)
        _check(_lib().gdtk_hip_mlp_loss_grad(self.h, X.shape[0], X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp),
                                             C.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def set_bf16_training(self, on=True):
        _check(_lib().gdtk_hip_mlp_set_bf16_training(self.h, 1 if on else 0))

    def train(self, n, x_ptr, y_ptr, batch=4096, epochs=1, lr=1e-3, seed=0):
        """Adam on device arrays x [n_in][n], y [n_out][n] (FP32 pointers); loss per epoch."""
        hist = np.zeros(epochs, dtype=np.float32)
        _check(_lib().gdtk_hip_mlp_train(self.h, n, x_ptr, y_ptr, batch, epochs, lr, seed, hist.ctypes.data_as(_fp),
                                         epochs))
        return hist

    def infer(self, n, x_ptr, y_ptr, stream=None):
        _check(_lib().gdtk_hip_mlp_infer(self.h, n, x_ptr, y_ptr, stream))


def air_element_matrix(species, M):
    """Element mass per unit species mass, rows N and O (each column sums to 1)."""
    atoms = element_matrix(species)
    m_el = np.array([0.0140067, 0.0159994])
    return atoms * m_el[:, None] / np.asarray(M)[None, :]


def features(rho, T, massf, dt):
    """Unscaled features (log T, log rho, Y[nsp], log dt) as [nsp + 3][n] float64 (numpy)."""
    n = rho.size
    Y = np.asarray(massf).reshape(-1, n)
    return np.vstack([np.log(T)[None, :], np.log(rho)[None, :], Y, np.full((1, n), np.log(dt))])


class Surrogate:
    """The learned step (gdtk_hip_surrogate_*) over a chemistry handle and two networks."""

    def __init__(self, chem, net_a, net_b, cfg: SurrogateCfg):
        self.chem, self.net_a, self.net_b, self.cfg = chem, net_a, net_b, cfg
        h = C.c_int(-1)
        _check(_lib().gdtk_hip_surrogate_create(chem.h, net_a.h, net_b.h, C.byref(cfg), C.byref(h)))
        self.h = h.value

    def close(self):
        if getattr(self, "h", -1) >= 0:
            _lib().gdtk_hip_surrogate_destroy(self.h)
            self.h = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step_device(self, n, rho, T, u, p, massf, dt_chem, dt, T_out, u_out, p_out, massf_out, dt_chem_out, status,
                    path=None, stream=None):
        """Device pointers (ints), asynchronous on stream.  path (optional int32 array): 0
        network, 1 integrator, 2 frozen, 3 failed."""
        _check(_lib().gdtk_hip_surrogate_step_device(self.h, n, rho, T, u, p, massf, dt_chem, dt, T_out, u_out,
                                                     p_out, massf_out, dt_chem_out, status, path, stream))

    def stats(self):
        """Of the last step: cells, network cells sent back to the integrator (fallback), the
        three phases' milliseconds, the largest network disagreement, network cells."""
        v = np.zeros(7)
        _check(_lib().gdtk_hip_surrogate_stats(self.h, v.ctypes.data_as(_dp), 7))
        return dict(cells=int(v[0]), fallback=int(v[1]), ms_network=v[2], ms_project=v[3], ms_integrator=v[4],
                    max_disagreement=v[5], network_cells=int(v[6]))


def project_device(E, n, Y0, d0, d, ok, stream=None):
    """The element-conservation projection alone (gdtk_hip_surrogate_project): E [n_elem][nsp]
    on the host; Y0, d0, d (float64 [nsp][n]) and ok (int32 [n]) are device pointers."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    _check(_lib().gdtk_hip_surrogate_project(E.shape[1], E.shape[0], E.ctypes.data_as(_dp), n, Y0, d0, d, ok,
                                             stream))


def integrator_samples(chem, cells, dt, device):
    """(state, dt, result) samples of the GPU alpha-QSS integrator for the given cells:
    returns device tensors of the inputs and the integrator's outputs."""
    import torch
    n = cells["rho"].size
    t = {k: torch.tensor(np.asarray(v), dtype=torch.float64, device=device) for k, v in cells.items()}
    dtc = torch.full((n,), -1.0, dtype=torch.float64, device=device)
    out = {k: torch.empty_like(t[k]) for k in ("T", "u", "p", "massf")}
    dto = torch.empty_like(dtc)
    st = torch.empty(n, dtype=torch.int32, device=device)
    chem.step_device(n, t["rho"].data_ptr(), t["T"].data_ptr(), t["u"].data_ptr(), t["p"].data_ptr(),
                     t["massf"].data_ptr(), dtc.data_ptr(), dt, out["T"].data_ptr(), out["u"].data_ptr(),
                     out["p"].data_ptr(), out["massf"].data_ptr(), dto.data_ptr(), st.data_ptr(),
                     torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return t, out, st


def train_air_surrogate(chem, gas_json, n_train=1 << 22, dt=1.0e-7, seed=7, **kw):
    """Samples the integrator on n_train cells of the air workload, trains two networks
    (different initialisations and shuffles) and returns (Surrogate, info)."""
    from .chemcases import air_cells
    return train_surrogate(chem, gas_json, air_cells(n_train, seed=seed), (dt,), seed=seed, **kw)


def train_surrogate(chem, gas_json, cells, dts, seed=7, epochs=(8, 6, 4), lrs=(2e-3, 5e-4, 1e-4), batch=4096,
                    x_abs_max=6.0, disagree_tol=5.0e-4, bf16_aware=False, log=None):
    """Trains two networks (different initialisations and shuffles) on the GPU integrator's
    results for the given cell states (dict rho, T, u, p, massf species-major) over each
    chemistry interval in dts, and returns (Surrogate, info)."""
    import torch

    dev = torch.device("cuda:0")
    nsp = chem.nsp
    t0 = time.perf_counter()
    feats, dYs = [], []
    for dt in dts:
        t_in, t_out, st = integrator_samples(chem, cells, dt, dev)
        ok = (st == 0) & (t_in["T"] > 300.0)
        Y0 = t_in["massf"].reshape(nsp, -1)
        dYs.append((t_out["massf"].reshape(nsp, -1) - Y0)[:, ok])
        feats.append(torch.vstack([torch.log(t_in["T"])[None, :], torch.log(t_in["rho"])[None, :], Y0,
                                   torch.full((1, Y0.shape[1]), float(np.log(dt)), dtype=torch.float64,
                                              device=dev)])[:, ok])
        del t_in, t_out, st
    f = torch.cat(feats, dim=1)
    dY = torch.cat(dYs, dim=1)
    del feats, dYs
    m = int(f.shape[1])
    x_mean = f.mean(dim=1)
    x_std = torch.clamp(f.std(dim=1), min=0.05)
    y_scale = torch.clamp(dY.std(dim=1), min=1e-6)
    X = ((f - x_mean[:, None]) / x_std[:, None]).to(torch.float32).contiguous()
    Yt = (dY / y_scale[:, None]).to(torch.float32).contiguous()
    t_samp = time.perf_counter() - t0
    nets, hist = [], []
    t1 = time.perf_counter()
    for k, s in enumerate((seed * 101 + 1, seed * 101 + 2)):
        net = Mlp(nsp + 3, nsp, seed=s)
        net.set_bf16_training(bf16_aware)
        h = []
        for ep, lr in zip(epochs, lrs):
            h += list(net.train(m, X.data_ptr(), Yt.data_ptr(), batch=batch, epochs=ep, lr=lr, seed=s + len(h)))
            if log:
                log(f"net {k}: lr {lr:g}, loss {h[-1]:.3e}")
        nets.append(net)
        hist.append(h)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    cfg = SurrogateCfg()
    for k in range(nsp + 3):
        cfg.x_mean[k] = float(x_mean[k])
        cfg.x_inv_std[k] = 1.0 / float(x_std[k])
    for k in range(nsp):
        cfg.y_scale[k] = float(y_scale[k])
    cfg.x_abs_max = x_abs_max
    cfg.disagree_tol = disagree_tol
    E = air_element_matrix(gas_json["species"], gas_json["M"])
    for e in range(E.shape[0]):
        for s in range(nsp):
            cfg.elem[e][s] = float(E[e, s])
    cfg.n_elem = E.shape[0]
    sur = Surrogate(chem, nets[0], nets[1], cfg)
    info = dict(samples=m, sample_s=t_samp, train_s=t_train, loss=hist, x_mean=x_mean.cpu().numpy(),
                x_std=x_std.cpu().numpy(), y_scale=y_scale.cpu().numpy())
    return sur, info


def with_tolerance(sur, disagree_tol, x_abs_max=None, route_ratio=None, route_T=None):
    """A second surrogate over the same networks and chemistry with other thresholds (and,
    if given, other routing: cells go to the networks where dt > route_ratio * dt_chem, or
    with no dt_chem yet where T > route_T; route_ratio <= 0 sends every cell above T_frozen)."""
    cfg = SurrogateCfg()
    C.memmove(C.byref(cfg), C.byref(sur.cfg), C.sizeof(SurrogateCfg))
    cfg.disagree_tol = disagree_tol
    if x_abs_max is not None:
        cfg.x_abs_max = x_abs_max
    if route_ratio is not None:
        cfg.route_ratio = route_ratio
    if route_T is not None:
        cfg.route_T = route_T
    return Surrogate(sur.chem, sur.net_a, sur.net_b, cfg)
