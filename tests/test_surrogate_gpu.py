"""GPU checks of the learned chemistry surrogate against the oracle (oracle/mlp_oracle.py)
and the alpha-QSS integrator.

- the matrix-core lane layout of the inference kernel (exact small-integer products);
- FP32 forward pass and gradients of the training kernels against the float64 oracle;
- BF16 inference against the oracle evaluated with the same BF16 roundings (inputs,
  weights and every activation rounded to BF16, FP32-or-better accumulation);
- training reduces the loss;
- the surrogate step: every network answer conserves each element's mass and keeps the
  mass fractions non-negative, frozen cells pass through, and cells handed to the
  integrator are bit-identical to the integrator's own results.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_mfma_probe_layout():
    from gdtk_amd.surrogate import mfma_probe
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float64)
    B = rng.integers(-4, 5, (32, 16)).astype(np.float64)
    D = mfma_probe(A, B)
    assert np.array_equal(D.astype(np.float64), A @ B)


@pytest.mark.parametrize("hidden,n_hidden", [(64, 2), (256, 3)])
def test_forward_and_gradient_match_oracle(hidden, n_hidden):
    from gdtk_amd.surrogate import Mlp
    net = Mlp(8, 5, hidden=hidden, n_hidden=n_hidden, seed=3)
    lay, npar = MO.layout(8, hidden, n_hidden, 5)
    assert net.nparams == npar
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(1)
    X = rng.normal(size=(77, 8))
    Y = rng.normal(size=(77, 5))
    y_ref, _, _ = MO.forward(p, lay, X)
    y_gpu = net.forward(X)
    assert np.max(np.abs(y_gpu - y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(y_ref)))
    loss_ref, g_ref = MO.loss_grad(p, lay, X, Y)
    loss, g = net.loss_grad(X, Y)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    for (i, o, w, b) in lay:
        for lo, hi in ((w, w + o * i), (b, b + o)):
            scale = np.max(np.abs(g_ref[lo:hi]))
            assert np.max(np.abs(g[lo:hi] - g_ref[lo:hi])) <= 1e-4 * scale, (lo, hi)


def _bf16_forward(p, lay, X):
    """The inference kernel's arithmetic: BF16 operands, exact products, wide sums."""
    Ws = MO.unpack(p, lay)
    h = _bf16(X)
    for l, (W, b) in enumerate(Ws):
        z = h @ _bf16(W).T + b.astype(np.float32).astype(np.float64)
        if l < len(Ws) - 1:
            h = _bf16(MO.gelu(z))
        else:
            h = z
    return h


def test_bf16_inference_matches_rounded_oracle():
    import torch
    from gdtk_amd.surrogate import Mlp
    net = Mlp(8, 5, seed=11)
    lay, _ = MO.layout(8, 256, 3, 5)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(2)
    n = 1000  # not a multiple of 64: exercises the partial last wave
    X = rng.normal(size=(n, 8)).astype(np.float32)
    dev = torch.device("cuda:0")
    xs = torch.tensor(np.ascontiguousarray(X.T), device=dev)
    ys = torch.empty((5, n), dtype=torch.float32, device=dev)
    net.infer(n, xs.data_ptr(), ys.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    y = ys.cpu().numpy().T.astype(np.float64)
    y_bf = _bf16_forward(p, lay, X.astype(np.float64))
    y_ref, _, _ = MO.forward(p, lay, X.astype(np.float64))
    scale = np.max(np.abs(y_ref))
    # same roundings: differences only from FP32 accumulation order and a rare BF16 tie
    assert np.max(np.abs(y - y_bf)) <= 2e-3 * scale, np.max(np.abs(y - y_bf)) / scale
    # and BF16 overall stays within a few per cent of the FP64 network
    assert np.max(np.abs(y - y_ref)) <= 5e-2 * scale


def test_bf16_aware_forward_matches_inference():
    """With BF16-aware training on, the training forward pass computes the inference
    kernel's function (BF16 operands and activations) up to FP32 accumulation order."""
    import torch
    from gdtk_amd.surrogate import Mlp
    net = Mlp(8, 5, seed=13)
    net.set_bf16_training(True)
    lay, _ = MO.layout(8, 256, 3, 5)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(3)
    n = 256
    X = rng.normal(size=(n, 8)).astype(np.float32)
    y_train = net.forward(X).astype(np.float64)
    dev = torch.device("cuda:0")
    xs = torch.tensor(np.ascontiguousarray(X.T), device=dev)
    ys = torch.empty((5, n), dtype=torch.float32, device=dev)
    net.infer(n, xs.data_ptr(), ys.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    y_inf = ys.cpu().numpy().T.astype(np.float64)
    y_bf = _bf16_forward(p, lay, X.astype(np.float64))
    scale = np.max(np.abs(y_bf))
    assert np.max(np.abs(y_train - y_bf)) <= 2e-3 * scale
    assert np.max(np.abs(y_inf - y_train)) <= 2e-3 * scale


def test_training_reduces_loss():
    import torch
    from gdtk_amd.surrogate import Mlp
    rng = np.random.default_rng(5)
    n = 1 << 15
    X = rng.normal(size=(8, n))
    Y = np.vstack([np.sin(X[0]), X[1] * X[2], np.tanh(X[3] + X[4]), 0.5 * X[5] ** 2 - 0.5, X[6] - X[7]])
    dev = torch.device("cuda:0")
    xs = torch.tensor(X, dtype=torch.float32, device=dev)
    ys = torch.tensor(Y, dtype=torch.float32, device=dev)
    net = Mlp(8, 5, seed=1)
    l0 = MO.loss_grad(net.get_params().astype(np.float64), MO.layout(8, 256, 3, 5)[0], X[:, :2048].T,
                      Y[:, :2048].T)[0]
    hist = net.train(n, xs.data_ptr(), ys.data_ptr(), batch=1024, epochs=4, lr=1e-3, seed=3)
    print("loss", l0, hist, flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.1 * l0


def test_surrogate_step_conservation_and_fallback():
    """A quickly trained surrogate on the air workload: structural guarantees (not accuracy,
    which the benchmark reports with the full training set)."""
    import torch
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, gas_record, mech_record
    from gdtk_amd.chemcases import air_cells
    from gdtk_amd.surrogate import air_element_matrix, integrator_samples, train_air_surrogate
    gd = gdata.load("tpg-air-5sp")
    ch = Chemistry(gas_record(gd), mech_record(gdata.load("park-air-5sp-1T")))
    # permissive thresholds: this checks the step's structure, so most cells should take
    # the network path even though two epochs of training leave the networks apart
    sur, info = train_air_surrogate(ch, gd, n_train=1 << 16, epochs=(2,), lrs=(2e-3,), batch=2048,
                                    x_abs_max=50.0, disagree_tol=1.0)
    n = 20000
    c = air_cells(n, seed=99)
    c["T"][:100] = np.linspace(200.0, 300.0, 100)  # frozen cells
    dev = torch.device("cuda:0")
    t_in, ref, ref_st = integrator_samples(ch, c, 1.0e-7, dev)
    out = {k: torch.empty_like(t_in[k]) for k in ("T", "u", "p", "massf")}
    dtc = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
    dto = torch.empty_like(dtc)
    st = torch.empty(n, dtype=torch.int32, device=dev)
    path = torch.empty(n, dtype=torch.int32, device=dev)
    sur.step_device(n, t_in["rho"].data_ptr(), t_in["T"].data_ptr(), t_in["u"].data_ptr(), t_in["p"].data_ptr(),
                    t_in["massf"].data_ptr(), dtc.data_ptr(), 1.0e-7, out["T"].data_ptr(), out["u"].data_ptr(),
                    out["p"].data_ptr(), out["massf"].data_ptr(), dto.data_ptr(), st.data_ptr(), path.data_ptr(),
                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    P = path.cpu().numpy()
    s = sur.stats()
    print("surrogate", s, flush=True)
    # default routing: every cell above T_frozen is offered to the networks
    assert s["network_cells"] == n - 100
    assert s["fallback"] == int(np.count_nonzero(P == 1))
    Y0 = c["massf"].reshape(5, n)
    Y = out["massf"].cpu().numpy().reshape(5, n)
    E = air_element_matrix(gd["species"], gd["M"])
    net_cells = P == 0
    assert net_cells.sum() > 0.5 * n
    dY = Y[:, net_cells] - Y0[:, net_cells]
    assert np.max(np.abs(E @ dY)) < 1e-14
    assert np.min(Y[:, net_cells]) >= 0.0
    assert np.max(np.abs(Y[:, net_cells].sum(axis=0) - 1.0)) < 1e-13
    # frozen cells pass through
    assert np.all(P[:100] == 2)
    for k in ("T", "u", "p"):
        assert np.array_equal(out[k].cpu().numpy()[:100], c[k][:100])
    # integrator cells equal the integrator's own results bit for bit
    fb = P == 1
    refY = ref["massf"].cpu().numpy().reshape(5, n)
    assert np.array_equal(Y[:, fb], refY[:, fb])
    assert np.array_equal(out["T"].cpu().numpy()[fb], ref["T"].cpu().numpy()[fb])
    assert np.array_equal(st.cpu().numpy()[fb], ref_st.cpu().numpy()[fb])
    err = np.max(np.abs(Y[:, net_cells] - refY[:, net_cells]))
    print("max |dY| error of network cells (two epochs of training)", err, flush=True)


def test_device_projection_matches_oracle():
    """The device's element-conservation projection (the one k_surr_apply runs on every
    network answer) against the oracle's active-set projection, which
    tests/test_surrogate_oracle.py pins to a general constrained least-squares solve, on the
    same (Y0, d0): clean cases, cases where one or more species hit zero (the active set
    grows), and infeasible ones (both refuse): an element that no species carries leaves
    E E^T singular."""
    import torch

    from gdtk_amd import data as gdata
    from gdtk_amd.surrogate import air_element_matrix, project_device
    species = ["N2", "O2", "NO", "N", "O"]
    E = air_element_matrix(species, gdata.load("tpg-air-5sp")["M"])
    rng = np.random.default_rng(11)
    n = 4096
    Y0 = rng.dirichlet(np.ones(5) * 0.6, size=n)
    tiny = rng.integers(0, 5, n)
    Y0[np.arange(n), tiny] *= 10.0 ** rng.uniform(-12, -2, n)
    Y0 /= Y0.sum(axis=1, keepdims=True)
    d0 = rng.normal(0.0, 1.0, (n, 5)) * (10.0 ** rng.uniform(-6, -0.5, (n, 1)))
    # a quarter of the cases: a pure O2 -> 2 O or N2 -> 2 N shift (already conserving)
    q = n // 4
    amt = rng.uniform(-0.2, 0.2, q)
    d0[:q] = 0.0
    d0[:q, 1], d0[:q, 4] = -amt, amt
    dev = torch.device("cuda:0")
    tY = torch.tensor(np.ascontiguousarray(Y0.T), dtype=torch.float64, device=dev)
    td0 = torch.tensor(np.ascontiguousarray(d0.T), dtype=torch.float64, device=dev)
    td = torch.empty_like(tY)
    tok = torch.empty(n, dtype=torch.int32, device=dev)

    def run(Em):
        project_device(Em, n, tY.data_ptr(), td0.data_ptr(), td.data_ptr(), tok.data_ptr(),
                       torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return td.cpu().numpy().T, tok.cpu().numpy()

    # a third element that no species carries: every cell is infeasible for both
    E3 = np.vstack([E, np.zeros((1, 5))])
    d_gpu, ok = run(E3)
    assert not ok.any() and not d_gpu.any()
    assert all(MO.project(Y0[i], d0[i], E3) is None for i in range(0, n, 64))
    d_gpu, ok = run(E)
    grew = worst = 0
    for i in range(n):
        ref = MO.project(Y0[i], d0[i], E)
        if ref is None:
            assert ok[i] == 0, i
            assert np.all(d_gpu[i] == 0.0)
            continue
        assert ok[i] == 1, i
        err = float(np.max(np.abs(d_gpu[i] - ref)))
        worst = max(worst, err)
        assert err <= 1e-13, (i, Y0[i], d0[i], d_gpu[i], ref)
        grew += int(np.any(Y0[i] + ref == 0.0))
    print(f"projection: {n} cells, {grew} with a species driven to zero, {int((ok == 0).sum())} infeasible, "
          f"max |d_gpu - d_oracle| {worst:.2e}", flush=True)
    assert grew > 100


def test_surrogate_routing_by_integration_length():
    """Routing: cells whose integration would be long go to the networks, the rest straight
    to the integrator.  With no dt_chem yet (-1) the test is T > route_T; with a dt_chem from
    a previous step it is dt > route_ratio * dt_chem.  Cells routed to the integrator equal
    the integrator's own results bit for bit, and every network cell conserves elements."""
    import torch
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, gas_record, mech_record
    from gdtk_amd.chemcases import air_cells
    from gdtk_amd.surrogate import integrator_samples, train_air_surrogate, with_tolerance
    gd = gdata.load("tpg-air-5sp")
    ch = Chemistry(gas_record(gd), mech_record(gdata.load("park-air-5sp-1T")))
    sur0, _ = train_air_surrogate(ch, gd, n_train=1 << 15, epochs=(1,), lrs=(2e-3,), batch=2048,
                                  x_abs_max=50.0, disagree_tol=1.0)
    n, dt = 8192, 1.0e-7
    c = air_cells(n, seed=7)
    dev = torch.device("cuda:0")
    t_in, ref, ref_st = integrator_samples(ch, c, dt, dev)
    out = {k: torch.empty_like(t_in[k]) for k in ("T", "u", "p", "massf")}
    dto = torch.empty(n, dtype=torch.float64, device=dev)
    st = torch.empty(n, dtype=torch.int32, device=dev)
    path = torch.empty(n, dtype=torch.int32, device=dev)
    # the integrator's own dt_chem suggestions of one step, as the next step would see them
    dtc_ref = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
    o2 = {k: torch.empty_like(t_in[k]) for k in ("T", "u", "p", "massf")}
    st2 = torch.empty_like(st)
    ch.step_device(n, t_in["rho"].data_ptr(), t_in["T"].data_ptr(), t_in["u"].data_ptr(), t_in["p"].data_ptr(),
                   t_in["massf"].data_ptr(), dtc_ref.data_ptr(), dt, o2["T"].data_ptr(), o2["u"].data_ptr(),
                   o2["p"].data_ptr(), o2["massf"].data_ptr(), dto.data_ptr(), st2.data_ptr(),
                   torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    dtc_prev = torch.where(st2 == 0, dto, torch.full_like(dto, -1.0))
    T = c["T"]
    for label, dtc_in, ratio, Tr, expect in (
            ("by T", torch.full((n,), -1.0, dtype=torch.float64, device=dev), 5.0, 6000.0, T > 6000.0),
            ("by dt_chem", dtc_prev, 5.0, 1.0e30, None)):
        sur = with_tolerance(sur0, 1.0, route_ratio=ratio, route_T=Tr)
        ref_i = {k: torch.empty_like(t_in[k]) for k in ("T", "u", "p", "massf")}
        dto_i, st_i = torch.empty_like(dto), torch.empty_like(st)
        ch.step_device(n, t_in["rho"].data_ptr(), t_in["T"].data_ptr(), t_in["u"].data_ptr(), t_in["p"].data_ptr(),
                       t_in["massf"].data_ptr(), dtc_in.data_ptr(), dt, ref_i["T"].data_ptr(), ref_i["u"].data_ptr(),
                       ref_i["p"].data_ptr(), ref_i["massf"].data_ptr(), dto_i.data_ptr(), st_i.data_ptr(),
                       torch.cuda.current_stream().cuda_stream)
        sur.step_device(n, t_in["rho"].data_ptr(), t_in["T"].data_ptr(), t_in["u"].data_ptr(), t_in["p"].data_ptr(),
                        t_in["massf"].data_ptr(), dtc_in.data_ptr(), dt, out["T"].data_ptr(), out["u"].data_ptr(),
                        out["p"].data_ptr(), out["massf"].data_ptr(), dto.data_ptr(), st.data_ptr(), path.data_ptr(),
                        torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        P = path.cpu().numpy()
        s = sur.stats()
        if expect is None:
            d = dtc_in.cpu().numpy()
            expect = np.where(d > 0.0, dt > ratio * d, T > Tr)
        offered = expect & (T > 300.0)
        print(f"routing {label}: {int(offered.sum())} cells offered to the networks, stats {s}", flush=True)
        assert s["network_cells"] == int(offered.sum()) and 0 < offered.sum() < n
        assert np.all(P[~offered & (T > 300.0)] == 1)
        assert np.all((P[offered] == 0) | (P[offered] == 1) | (P[offered] == 3))
        fb = P == 1
        for k in ("T", "u", "p"):
            assert np.array_equal(out[k].cpu().numpy()[fb], ref_i[k].cpu().numpy()[fb]), (label, k)
        assert np.array_equal(out["massf"].cpu().numpy().reshape(5, n)[:, fb],
                              ref_i["massf"].cpu().numpy().reshape(5, n)[:, fb])
        assert np.array_equal(dto.cpu().numpy()[fb], dto_i.cpu().numpy()[fb])
        sur.close()


def test_trained_network_gains_over_the_integrator():
    """What the trained networks gain over the method without them, on a held-out batch of
    2^18 air cells of the chemistry workload (T 1500-12000 K, dt = 1e-7 s, uneven
    stiffness; networks trained on 2^20 other cells of the same distribution, 13 epochs):
    - the cells the networks answer (most of them) take a fraction of the time the alpha-QSS
      integrator takes over those same cells, and stay within the stated bounds of the
      integrator's result;
    - the whole step (networks, then the integrator for the cells they send back) is timed
      beside the integrator's whole step.  It gains nothing here: the cells sent back are
      among the stiffest, and one such cell's serial substeps set the integrator's time.
    The full-size figures are the benchmark's."""
    import torch
    from gdtk_amd import data as gdata
    from gdtk_amd.chem import Chemistry, gas_record, mech_record
    from gdtk_amd.chemcases import air_cells
    from gdtk_amd.surrogate import integrator_samples, train_air_surrogate, with_tolerance
    gd = gdata.load("tpg-air-5sp")
    ch = Chemistry(gas_record(gd), mech_record(gdata.load("park-air-5sp-1T")))
    dt = 1.0e-7
    sur0, info = train_air_surrogate(ch, gd, n_train=1 << 20, dt=dt, seed=3, epochs=(6, 4, 3),
                                     lrs=(2e-3, 5e-4, 1e-4), batch=4096)
    sur = with_tolerance(sur0, 5.0e-3)
    n = 1 << 18
    c = air_cells(n, seed=424242)
    dev = torch.device("cuda:0")
    t_in, ref, ref_st = integrator_samples(ch, c, dt, dev)
    stream = torch.cuda.current_stream()

    def run(f, reps=3):
        f()  # warm-up
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            f()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def bufs(m):
        return ({k: torch.empty((5 * m,) if k == "massf" else (m,), dtype=torch.float64, device=dev)
                 for k in ("T", "u", "p", "massf", "dtc")}, torch.empty(m, dtype=torch.int32, device=dev))

    def args(t, m, o, st):
        dtc = torch.full((m,), -1.0, dtype=torch.float64, device=dev)
        bufs_keep.append(dtc)
        return (m, t["rho"].data_ptr(), t["T"].data_ptr(), t["u"].data_ptr(), t["p"].data_ptr(), t["massf"].data_ptr(),
                dtc.data_ptr(), dt, o["T"].data_ptr(), o["u"].data_ptr(), o["p"].data_ptr(), o["massf"].data_ptr(),
                o["dtc"].data_ptr(), st.data_ptr())

    bufs_keep = []
    o, st = bufs(n)
    path = torch.empty(n, dtype=torch.int32, device=dev)
    a = args(t_in, n, o, st)
    ms_int_all = run(lambda: ch.step_device(*a, stream.cuda_stream))
    ms_sur_all = run(lambda: sur.step_device(*a, path.data_ptr(), stream.cuda_stream))
    s = sur.stats()
    P = path.cpu().numpy()
    ok = (P == 0) & (ref_st.cpu().numpy() == 0)
    # the network-answered cells alone, through each path
    idx = torch.tensor(np.nonzero(P == 0)[0], dtype=torch.int64, device=dev)
    m = int(idx.numel())
    sub = {"rho": t_in["rho"][idx].contiguous(), "T": t_in["T"][idx].contiguous(), "u": t_in["u"][idx].contiguous(),
           "p": t_in["p"][idx].contiguous(),
           "massf": t_in["massf"].reshape(5, n)[:, idx].contiguous().reshape(-1)}
    o2, st2 = bufs(m)
    path2 = torch.empty(m, dtype=torch.int32, device=dev)
    a2 = args(sub, m, o2, st2)
    ms_int_sub = run(lambda: ch.step_device(*a2, stream.cuda_stream))
    ms_net_sub = run(lambda: sur.step_device(*a2, path2.data_ptr(), stream.cuda_stream))
    sent_back = int(sur.stats()["fallback"])
    Y = o["massf"].cpu().numpy().reshape(5, n)
    Yr = ref["massf"].cpu().numpy().reshape(5, n)
    dY = np.abs(Y - Yr).max(axis=0)[ok]
    Tr = ref["T"].cpu().numpy()
    Terr = (np.abs(o["T"].cpu().numpy() - Tr) / Tr)[ok]
    rep = dict(network_fraction=float(ok.mean()), answered_cells=m, integrator_ms_on_answered=ms_int_sub,
               networks_ms_on_answered=ms_net_sub, gain_on_answered=ms_int_sub / ms_net_sub,
               whole_step_integrator_ms=ms_int_all, whole_step_surrogate_ms=ms_sur_all,
               whole_step_sent_back=s["fallback"], max_dY=float(dY.max()), p99_dY=float(np.quantile(dY, 0.99)),
               p99_T=float(np.quantile(Terr, 0.99)), samples=info["samples"], train_s=info["train_s"])
    print("surrogate gain:", rep, flush=True)
    assert rep["network_fraction"] > 0.9
    assert sent_back <= 0.001 * m       # the answered cells are answered again
    assert rep["gain_on_answered"] > 3.0
    assert rep["p99_dY"] < 1.0e-2 and rep["max_dY"] < 0.1 and rep["p99_T"] < 0.05
    sur.close()
