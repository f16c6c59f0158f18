"""Learned surrogate: BF16 MFMA network, training, and the guarded reactor path.

* Loss and gradient of the GPU network against a float64 numpy restatement of
  the same network that rounds to BF16 exactly where the GPU does (features,
  weights, activations, gradient operands): this pins the three GEMM operand
  layouts (forward, weight gradient, input gradient) and the MFMA lane layout.
  Bars: loss relative error <= 1e-4; gradient relative L2 error <= 2e-2 per
  layer (an activation that lands on a BF16 rounding boundary may round the
  other way, the only source of differences beyond FP32 summation order).
* Training on integrator pairs lowers the loss.
* The guarded advance keeps every reactor within 1 K of the plain FP64 BDF
  result, and an untrained network is rejected by the guard so the results
  are those of the integrator.
"""
import numpy as np
import pytest

from conftest import mech_file

pytestmark = pytest.mark.gpu


def bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


C_G, A_G = np.float32(0.7978845608028654), np.float32(0.044715)


def gelu(z):
    t = np.tanh(C_G * (z + A_G * z * z * z))
    return np.float32(0.5) * z * (np.float32(1) + t)


def gelu_grad(z):
    t = np.tanh(C_G * (z + A_G * z * z * z))
    return (np.float32(0.5) * (1 + t) + np.float32(0.5) * z * (1 - t * t) * C_G * (1 + 3 * A_G * z * z)).astype(np.float32)


def pad(x):
    return (x + 63) // 64 * 64


def reference_loss_grad(params, norm, K, hidden, depth, y0, P, y1):
    n_in, n_out = K + 2, K + 1
    ip, op = pad(n_in), pad(n_out)
    fi = [ip] + [hidden] * depth
    fo = [hidden] * depth + [op]
    W, b, off = [], [], 0
    for l in range(depth + 1):
        W.append(params[off:off + fo[l] * fi[l]].reshape(fo[l], fi[l]))
        off += fo[l] * fi[l]
        b.append(params[off:off + fo[l]])
        off += fo[l]
    Wb = [bf16(w).astype(np.float64) for w in W]
    mu_i, sd_i = norm[:ip], norm[ip:2 * ip]
    mu_o, sd_o, sc_o = norm[2 * ip:2 * ip + op], norm[2 * ip + op:2 * ip + 2 * op], norm[2 * ip + 2 * op:]
    rows = len(y0)
    f = np.zeros((rows, ip))
    f[:, 0] = y0[:, 1]
    f[:, 1] = np.log(P)
    f[:, 2:n_in] = np.log(np.maximum(y0[:, 2:], 1e-30))
    z = np.clip((f - mu_i) / sd_i, -8, 8)
    z[:, n_in:] = 0
    X = bf16(z.astype(np.float32)).astype(np.float64)
    T = np.zeros((rows, op), dtype=np.float32)
    T[:, :n_out] = ((np.arcsinh((y1[:, 1:] - y0[:, 1:]) / sc_o[:n_out]) - mu_o[:n_out]) / sd_o[:n_out]).astype(np.float32)
    H, Z = [X], []
    for l in range(depth):
        z = (H[-1] @ Wb[l].T).astype(np.float32) + b[l]
        Z.append(z)
        H.append(bf16(gelu(z)).astype(np.float64))
    out = (H[-1] @ Wb[depth].T).astype(np.float32) + b[depth]
    d = (out - T)[:, :n_out].astype(np.float64)
    scale = 1.0 / (rows * n_out)
    loss = float((d * d).sum() * scale)
    G = np.zeros((rows, op), dtype=np.float32)
    G[:, :n_out] = np.float32(2 * scale) * (out - T)[:, :n_out]
    G = bf16(G).astype(np.float64)
    grads = [None] * (depth + 1)
    for l in range(depth, -1, -1):
        dW = (G.T @ H[l]).astype(np.float32)
        db = G.sum(axis=0).astype(np.float32)
        grads[l] = (dW, db)
        if l > 0:
            G = bf16((G @ Wb[l]).astype(np.float32) * gelu_grad(Z[l - 1])).astype(np.float64)
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat, [fo[l] * fi[l] + fo[l] for l in range(depth + 1)]


def synthetic_pairs(info, n, seed):
    g = np.random.default_rng(seed)
    K = info.n_species
    X = 10.0 ** g.uniform(-8, 0, (n, K))
    X /= X.sum(axis=1, keepdims=True)
    Y = info.mass_fractions_from_moles(X)
    y0 = np.concatenate([np.ones((n, 1)), g.uniform(900, 2200, (n, 1)), Y], axis=1)
    P = 101325.0 * g.uniform(1, 20, n)
    y1 = y0.copy()
    y1[:, 1] += g.normal(0, 5.0, n)
    y1[:, 2:] += 1e-4 * g.normal(0, 1, (n, K)) * Y
    return y0, P, y1


@pytest.mark.parametrize("mech,hidden,depth,n", [("h2o2", 64, 2, 300), ("gri30", 256, 4, 1000)])
def test_loss_and_gradient_match_reference(mech, hidden, depth, n):
    import cantera_amd as ca
    from cantera_amd.surrogate import Surrogate
    from cantera_amd.workloads import MechInfo
    info = MechInfo(mech_file(mech))
    sol = ca.Solution(mech)
    sur = Surrogate(sol, dt=1e-6, hidden=hidden, depth=depth, seed=3)
    y0, P, y1 = synthetic_pairs(info, n, 5)
    sur.fit_normalization(y0, P, y1)
    loss, g = sur.loss_grad(y0, P, y1)
    rl, rg, sizes = reference_loss_grad(sur.params, sur.normalization, sol.n_species, hidden, depth, y0, P, y1)
    assert abs(loss - rl) <= 1e-4 * rl
    off = 0
    for l, sz in enumerate(sizes):
        a, b = g[off:off + sz].astype(np.float64), rg[off:off + sz].astype(np.float64)
        err = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        print(f"{mech} layer {l}: gradient relative error {err:.2e}", flush=True)
        assert err <= 2e-2
        off += sz


def test_training_lowers_loss():
    import cantera_amd as ca
    from cantera_amd.surrogate import Surrogate, generate_pairs
    from cantera_amd.workloads import MechInfo, config1
    info = MechInfo(mech_file("h2o2"))
    sol = ca.Solution("h2o2")
    wl = config1(info, n=256)
    y0, P, y1 = generate_pairs(sol, wl.T0, wl.P, wl.Y0, dt=1e-6, nsteps=60)
    sur = Surrogate(sol, dt=1e-6, hidden=128, depth=3, seed=1)
    losses = sur.fit(y0, P, y1, epochs=25, batch=512, lr=1e-3)
    print("losses", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.2 * losses[0]


def _plain_and_surrogate(sol, wl, sur, nsteps, dt, precision=0):
    """The same reactors advanced nsteps outputs of dt by the plain integrator
    and by the guarded surrogate path; returns both trajectories (n x steps+1
    x N), the counts [eligible, kept, integrated] and the two wall times."""
    import time

    import cantera_amd as ca
    a = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    b = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=precision)
    n = len(wl.T0)
    ya = np.zeros((n, nsteps + 1, sol.n_species + 2))
    yb = np.zeros_like(ya)
    ya[:, 0] = a.state
    yb[:, 0] = b.state
    tot = np.zeros(3, dtype=np.int64)
    ta = tb = 0.0
    for k in range(1, nsteps + 1):
        t0 = time.perf_counter()
        a.advance(k * dt)
        ta += time.perf_counter() - t0
        t0 = time.perf_counter()
        tot += b.advance_surrogate(k * dt, sur)
        tb += time.perf_counter() - t0
        ya[:, k] = a.state
        yb[:, k] = b.state
    return ya, yb, tot, ta, tb


def test_guarded_advance_h2o2_within_budget():
    """h2o2 at dt = 1e-6 s with a network trained on gated integrator pairs:
    whatever the guard keeps, every output stays within 1 K and 1e-4 in Y of
    the plain integrator, and mass fractions keep summing to 1."""
    import cantera_amd as ca
    from cantera_amd.surrogate import Surrogate, generate_pairs
    from cantera_amd.workloads import MechInfo, config1
    info = MechInfo(mech_file("h2o2"))
    sol = ca.Solution("h2o2")
    train = config1(info, n=512)
    y0, P, y1 = generate_pairs(sol, train.T0, train.P, train.Y0, dt=1e-6, nsteps=80, gate_T=1500.0)
    sur = Surrogate(sol, dt=1e-6, hidden=128, depth=3, seed=2)
    sur.fit(y0, P, y1, epochs=30, batch=512, lr=1e-3)
    wl = config1(info, n=96)
    ya, yb, tot, _, _ = _plain_and_surrogate(sol, wl, sur, 40, 1e-6)
    dT = np.abs(ya[:, :, 1] - yb[:, :, 1]).max()
    dY = np.abs(ya[:, :, 2:] - yb[:, :, 2:]).max()
    print(f"h2o2: {len(y0)} gated training pairs; eligible {tot[0]}, kept {tot[1]}, integrated {tot[2]}; "
          f"max |dT| {dT:.3e} K, max |dY| {dY:.3e}", flush=True)
    assert dT <= 1.0 and dY <= 1e-4
    assert np.all(np.abs(yb[:, -1, 2:].sum(axis=1) - 1) < 1e-5)


def test_trained_gri30_surrogate_gain_and_budget():
    """gri30 at the fixed step dt = 1e-5 s: a network trained on gated pairs
    of FP64 GPU trajectories (8192 reactors of the headline distribution,
    another seed, 100 steps) keeps guarded steps on 512 evaluation reactors
    over 100 steps, where an untrained network keeps none; at every output the
    guarded trajectories stay within the budget of 1 K in T and 1e-4 in Y of
    the plain FP64 integrator.  The wall-time ratio of the two paths is
    printed as measured."""
    import cantera_amd as ca
    from cantera_amd.surrogate import Surrogate, generate_pairs
    from cantera_amd.workloads import MechInfo, config3
    info = MechInfo(mech_file("gri30"))
    sol = ca.Solution("gri30")
    dt = 1e-5
    tr = config3(info, n=8192, nsteps=100, seed=0x5EED0104)
    y0, P, y1 = generate_pairs(sol, tr.T0, tr.P, tr.Y0, dt=dt, nsteps=100, rtol=tr.rtol, atol=tr.atol,
                               gate_T=ca.ReactorBatch.GUARD["gate_T"])
    sur = Surrogate(sol, dt=dt, hidden=256, depth=4, seed=1)
    losses = sur.fit(y0, P, y1, epochs=20, batch=4096, lr=1e-3)
    wl = config3(info, n=512, nsteps=100)
    ya, yb, tot, ta, tb = _plain_and_surrogate(sol, wl, sur, 100, dt)
    dT = np.abs(ya[:, :, 1] - yb[:, :, 1]).max()
    dY = np.abs(ya[:, :, 2:] - yb[:, :, 2:]).max()
    fresh = Surrogate(sol, dt=dt, hidden=256, depth=4, seed=5)
    fresh.fit_normalization(y0, P, y1)
    _, _, tot0, _, _ = _plain_and_surrogate(sol, config3(info, n=512, nsteps=20), fresh, 20, dt)
    print(f"gri30 dt 1e-5: {len(y0)} gated pairs, loss {losses[0]:.3g} -> {losses[-1]:.3g}; trained: eligible "
          f"{tot[0]}, kept {tot[1]} ({tot[1] / max(1, tot[0]):.3%}), integrated {tot[2]}; untrained (20 steps): kept "
          f"{tot0[1]}; max |dT| {dT:.3e} K, max |dY| {dY:.3e}; wall time plain {ta:.2f} s, guarded {tb:.2f} s "
          f"(gain {ta / tb:.3f}x)", flush=True)
    assert tot[1] > 0
    assert tot0[1] == 0
    assert dT <= 1.0 and dY <= 1e-4


def test_untrained_network_falls_back_to_integrator():
    import cantera_amd as ca
    from cantera_amd.surrogate import Surrogate
    from cantera_amd.workloads import MechInfo, config3
    info = MechInfo(mech_file("gri30"))
    sol = ca.Solution("gri30")
    wl = config3(info, n=64, nsteps=5)
    sur = Surrogate(sol, dt=1e-6, hidden=256, depth=4, seed=9)
    ya, yb, tot, _, _ = _plain_and_surrogate(sol, wl, sur, 5, 1e-6)
    print("counts", tot, flush=True)
    assert tot[1] == 0 and tot[2] == tot[0] == 5 * 64
    assert np.max(np.abs(ya[:, -1, 1] - yb[:, -1, 1]) / ya[:, -1, 1]) <= 1e-6


def test_prediction_conserves_elements():
    """predict() projects the mass-fraction increments onto the element-
    conserving subspace (one wavefront per row, Gram matrix by wavefront sums):
    element totals sum_k a_ek Y_k / W_k and sum_k Y_k of every predicted state
    equal those of its input to rounding, whatever the network outputs (here an
    untrained network on a fitted normalisation)."""
    import cantera_amd as ca
    from cantera_amd.surrogate import Surrogate
    from cantera_amd.workloads import MechInfo
    for mech in ("h2o2", "gri30", "nDodecane"):
        info = MechInfo(mech_file(mech))
        sol = ca.Solution(mech)
        sur = Surrogate(sol, dt=1e-6, hidden=64, depth=2, seed=4)
        y0, P, y1 = synthetic_pairs(info, 500, 6)
        sur.fit_normalization(y0, P, y1)
        yp = sur.predict(y0, P)
        A = info.composition / info.molecular_weights[:, None]     # K x E
        e0, e1 = y0[:, 2:] @ A, yp[:, 2:] @ A
        scale = (np.abs(y0[:, 2:]) + np.abs(yp[:, 2:])) @ np.abs(A) + 1e-300
        err = float(np.max(np.abs(e1 - e0) / scale))
        dsum = float(np.max(np.abs(yp[:, 2:].sum(axis=1) - y0[:, 2:].sum(axis=1))
                            / (np.abs(y0[:, 2:]).sum(axis=1) + np.abs(yp[:, 2:]).sum(axis=1))))
        print(f"{mech}: element error {err:.2e}, sum Y change {dsum:.2e}", flush=True)
        assert err <= 1e-10
        assert dsum <= 1e-10
        assert np.all(yp[:, 0] == y0[:, 0])
