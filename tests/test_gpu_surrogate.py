"""The microphysics surrogate in the model step (no reference counterpart).

On the 128 x 128 x 40 Thompson workload after 150 steps: teacher pairs from 180 more plain steps
(every point holding condensate), GPU training (12000 Adam steps of 8192 rows), then
  - the BF16 MFMA inference form against the oracle's quantised forward pass (float64 sums of
    the same BF16 values): relative L2 <= 2e-3 (FP32 accumulation order and BF16 rounding of
    activations that land on either side of a rounding boundary);
  - routing: with an envelope that admits nothing the step equals the plain step (to FP32
    rounding: the two physics-kernel variants may contract multiply-adds differently);
    with a network that predicts NaN every handed point falls back to the plain block;
  - the judge on one step, and what the trained network gains over the step without it (the
    handed points' microphysics left out, i.e. a network predicting zero tendencies);
  - 20 surrogate steps against 20 plain steps from the same state: column condensate.
"""
import os
import sys

import numpy as np
import pytest

from _util import rel_l2

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def snapshot(m):
    return {"V": [m.V[0].download(), m.V[1].download()], "Q1": [m.Q1[0].download(), m.Q1[1].download()],
            "Q2": [m.Q2[0].download(), m.Q2[1].download()], "meta": (m.time, m.lsfT, m.parity, m.steps)}


def restore(m, snap):
    for c in (0, 1):
        m.V[c].upload(snap["V"][c])
        m.Q1[c].upload(snap["Q1"][c])
        m.Q2[c].upload(snap["Q2"][c])
    m.time, m.lsfT, m.parity, m.steps = snap["meta"]


def column_condensate(m, Q1, Q2):
    rho = np.asarray(m._B, np.float64)[2:8 * m.nz:8][: m.nz]
    c = (Q1[..., 1] + Q1[..., 2] + Q1[..., 3] + Q2[..., 0] + Q2[..., 1]).astype(np.float64)
    return np.tensordot(rho, c, axes=(0, 0)) * m.dz


@pytest.fixture(scope="module")
def trained():
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_128
    from outbreaklab_amd.surrogate import Surrogate, fit, teach

    m = RegionalModel()
    dt = setup_plain_128(m, scheme="thompson")
    for _ in range(150):
        m.step(dt)
    rows = teach(m, dt, 180, capacity=2_000_000, p_active=1.0)
    net = Surrogate(seed=3)
    losses = fit(net, rows, steps=12000, batch=8192, lr=2e-3)
    print(f"teacher rows {rows.shape[0]}, training loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    yield m, dt, net, rows
    m.set_surrogate(None)
    net.destroy()
    m.destroy()


def test_inference_form_matches_oracle(trained):
    from outbreaklab_amd.surrogate import unpack

    _, _, net, rows = trained
    x = rows[:4096]
    y = net.infer(x)
    P = unpack(net.params())
    ref = M.forward_bf16(P, M.features(x[:, :14].astype(np.float64)))
    e = rel_l2(y[:, :9], ref[:, :9])
    _, _, yq = net.eval(x, quant=True)
    e2 = rel_l2(y[:, :9], yq[:, :9])
    print(f"inference vs oracle bf16 forward {e:.2e}, vs training forward {e2:.2e}")
    assert e <= 2e-3 and e2 <= 2e-3


def test_empty_envelope_is_the_plain_step(trained):
    from outbreaklab_amd.surrogate import Surrogate

    m, dt, net, _ = trained
    snap = snapshot(m)
    m.set_surrogate(None)
    m.step(dt)
    plain = m.state()
    restore(m, snap)
    closed = Surrogate(seed=1)
    closed.set_scales(np.ones(9, np.float32), np.full(14, 3e38, np.float32), np.full(14, -3e38, np.float32))
    m.set_surrogate(closed)
    m.step(dt)
    st = m.surrogate_stats()
    got = m.state()
    m.set_surrogate(None)
    closed.destroy()
    restore(m, snap)
    assert st["points"] == 0
    errs = [rel_l2(a, b) for a, b in zip(got, plain)]
    print("empty envelope vs plain:", errs)
    for e in errs:
        assert e <= 1e-6


def test_nan_predictions_fall_back_to_plain(trained):
    from outbreaklab_amd.surrogate import Surrogate

    m, dt, net, _ = trained
    snap = snapshot(m)
    m.set_surrogate(None)
    m.step(dt)
    plain = m.state()
    restore(m, snap)
    bad = Surrogate(seed=1)
    p = bad.params()
    p[:] = np.nan
    bad.set_params(p)
    bad.set_scales(np.ones(9, np.float32))
    m.set_surrogate(bad)
    m.step(dt)
    st = m.surrogate_stats()
    got = m.state()
    m.set_surrogate(None)
    bad.destroy()
    restore(m, snap)
    print(f"handed {st['points']:.0f} points, fell back {st['fallbacks']:.0f}")
    assert st["points"] > 1000 and st["fallbacks"] == st["points"]
    for a, b in zip(got, plain):
        assert rel_l2(a, b) <= 1e-6


def test_judge_and_gain_over_no_surrogate(trained):
    from outbreaklab_amd.surrogate import Surrogate

    m, dt, net, _ = trained
    snap = snapshot(m)
    res = {}
    zero = Surrogate(seed=1)
    zero.set_params(np.zeros(zero.n, np.float32))
    zero.set_scales(np.ones(9, np.float32), net.in_lo, net.in_hi)
    for name, n in (("trained", net), ("zero", zero)):
        restore(m, snap)
        m.set_surrogate(n)
        m.step(dt, judge=True)
        res[name] = m.surrogate_stats(reset=True)
        m.set_surrogate(None)
    zero.destroy()
    restore(m, snap)
    t, z = res["trained"], res["zero"]
    print("trained:", {k: round(v, 6) for k, v in t["rel_l2"].items()}, "cc", t["column_condensate_rel_l2"],
          "dBZ rms", t["dbz_rms"], "precip", t["precip_rel_l2"])
    print("zero:   ", {k: round(v, 6) for k, v in z["rel_l2"].items()}, "cc", z["column_condensate_rel_l2"],
          "dBZ rms", z["dbz_rms"], "precip", z["precip_rel_l2"])
    assert t["judged"] == 1 and t["points"] > 1000 and t["fallbacks"] == 0
    # the trained network must recover most of the microphysics the step would miss without it
    for ch in ("qv", "qc", "qr", "theta"):
        assert t["rel_l2"][ch] < 0.5 * z["rel_l2"][ch], ch
    assert t["column_condensate_rel_l2"] < 0.5 * z["column_condensate_rel_l2"]
    assert t["column_condensate_rel_l2"] <= 0.05


def test_multistep_column_condensate(trained):
    m, dt, net, _ = trained
    snap = snapshot(m)
    m.set_surrogate(None)
    for _ in range(20):
        m.step(dt)
    _, q1p, q2p = m.state()
    restore(m, snap)
    m.set_surrogate(net)
    for _ in range(20):
        m.step(dt)
    _, q1s, q2s = m.state()
    st = m.surrogate_stats(reset=True)
    m.set_surrogate(None)
    restore(m, snap)
    ccp, ccs = column_condensate(m, q1p, q2p), column_condensate(m, q1s, q2s)
    e = rel_l2(ccs, ccp)
    print(f"20 steps: column condensate rel L2 {e:.4f} (total {ccs.sum():.4g} vs {ccp.sum():.4g}), "
          f"handed {st['points'] / 20:.0f} points/step")
    assert e <= 0.05
