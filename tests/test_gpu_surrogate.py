"""What the learned surrogate gains, and that it changes nothing else.

A network trained on the GPU (features and exact per-molecule energies of
MD configurations at 240-300 K and several densities) screens NPT Monte
Carlo proposals.  The exact FP64 energy still makes every final decision, so
(1) the chain with the network samples the same ensemble as plain MC (equal
<V> and <U> within the errors of independent replica chains), and (2) it needs far fewer exact
energy evaluations per accepted move than plain MC or the same scheme with
an untrained (zero) network."""
import numpy as np
import pytest

from genesis6_amd import g6, surrogate as S

pytestmark = pytest.mark.gpu

T, PRESS = 250.0, 1.0


def box(n, rho=1.0):
    return (n * 18.015 / (6.0221e23 * rho) * 1e24) ** (1.0 / 3.0)


@pytest.fixture(scope="module")
def trained():
    n = 512
    X, E = S.md_dataset(n, box(n), [240.0, 270.0, 300.0], 60, 20, 4242, rc=8.0,
                        scales=(0.985, 1.0, 1.015))
    net, losses = S.train(X, E, epochs=20, batch=4096, lr=2e-3)
    print("trained on %d rows, loss %.3f -> %.3f" % (len(X), losses[0], losses[-1]), flush=True)
    assert losses[-1] < 0.6 * losses[0] or losses[-1] < 0.5
    return net


def npt_system(n, seed, rc=8.0):
    L = box(n)
    s = g6.System(g6.TIP4P, n, L, rc, 0.5, 0.002, prec=g6.FP32, skin=1.0)
    com, q = g6.lattice_start(n, L, seed)
    s.upload(com, q)
    s.set_nose(T, T * 2500.0)
    s.vreset(T, seed)
    s.step(1500)
    return s


def test_trained_network_cuts_exact_evaluations(trained):
    pv = 1.0 / 28.0
    s = npt_system(512, 7)
    s.npt_begin(T, PRESS, 11, 0.15, 8.0, 0.005, pv, 8.0)
    s.npt_run(500)
    a = s.npt_run(0)
    b = s.npt_run(3000)
    plain = (b["proposals"] - a["proposals"]) / max(1, b["accepted"] - a["accepted"])
    res = {}
    for name, net in (("trained", trained), ("zero", g6.Mlp(np.zeros(g6.MLP_NPARAM, np.float32)))):
        s.npt_begin(T, PRESS, 11, 0.15, 8.0, 0.005, pv, 8.0)
        s.npt_run(500)
        a = s.npt_run(0)
        b = s.npt_run(300, mlp=net, K=32)
        res[name] = (b["exact_evaluations"] - a["exact_evaluations"]) / max(
            1, b["accepted"] - a["accepted"])
        print(name, {k: b[k] - a[k] for k in b}, flush=True)
    print("exact evaluations per accepted move: plain %.1f, zero network %.1f, trained %.1f"
          % (plain, res["zero"], res["trained"]), flush=True)
    assert res["trained"] * 4 < plain
    assert res["trained"] * 4 < res["zero"]


def _series(s, nblocks, per, **kw):
    v, u = [], []
    for _ in range(nblocks):
        s.npt_run(per, **kw)
        _, _, L, U = s.npt_state()
        v.append(L ** 3)
        u.append(U)
    return np.array(v), np.array(u)


def test_ensemble_averages_agree(trained):
    """<V> and <U> of plain MC and of delayed acceptance agree.  From one
    equilibrated 300 K state, six plain and six delayed-acceptance chains
    run with their own seeds; the first fifth of each is dropped and the
    means of the chain means are compared with the standard error from the
    spread between chains (4.5 combined errors).  Independent replicas give
    honest errors where the volume decorrelates slowly; blocks of one chain
    do not."""
    n, pv, voxa, temp = 256, 0.1, 4.5, 300.0
    L = box(n)
    s = g6.System(g6.TIP4P, n, L, 8.0, 0.5, 0.002, prec=g6.FP32, skin=1.0)
    com, q = g6.lattice_start(n, L, 21)
    s.upload(com, q)
    s.set_nose(temp, temp * 2500.0)
    s.vreset(temp, 21)
    s.step(1500)
    s.npt_begin(temp, PRESS, 31, 0.15, 8.0, 0.005, pv, voxa)
    s.npt_run(120000)
    com, q, L, _ = s.npt_state()
    means = {"plain": [], "delayed acceptance": []}
    for r in range(6):
        for name in means:
            c = g6.System(g6.TIP4P, n, L, 8.0, 0.5, 0.002, prec=g6.FP32, skin=1.0)
            c.upload(com, q.T.copy())
            c.npt_begin(temp, PRESS, 100 + 2 * r + (name != "plain"), 0.15, 8.0, 0.005, pv, voxa)
            if name == "plain":
                v, u = _series(c, 300, 250)
            else:
                v, u = _series(c, 300, 120, mlp=trained, K=4)
            means[name].append((v[60:].mean(), u[60:].mean()))
            print(name, r, "<V> %.1f <U> %.1f" % means[name][-1], flush=True)
    a, b = np.array(means["plain"]), np.array(means["delayed acceptance"])
    for col, q_name in ((0, "V"), (1, "U")):
        m1, m2 = a[:, col].mean(), b[:, col].mean()
        e1 = a[:, col].std(ddof=1) / np.sqrt(len(a))
        e2 = b[:, col].std(ddof=1) / np.sqrt(len(b))
        print("<%s> plain %.6g +- %.2g, delayed acceptance %.6g +- %.2g" % (q_name, m1, e1, m2, e2),
              flush=True)
        assert abs(m1 - m2) < 4.5 * np.hypot(e1, e2)
