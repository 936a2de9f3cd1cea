"""GPU tests of the learned time-of-impact proposer: the HIP network against
the CPU network (oracle/mlp.py, itself pinned by finite differences), the
dataset labels against the coplanarity cubic, training, determinism, and the
learned form's results against the CPU oracle bit for bit."""
import numpy as np
import pytest

from collision_amd import scenes as S
from collision_amd.solver import CollisionSolver3d, solver_for_scene
from oracle import mlp as M
from tests.parity import run_parity

pytestmark = pytest.mark.gpu


def _weights(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in M.SHAPES:
        if len(shape) == 2:
            lim = np.sqrt(6.0 / shape[1])
            d[name] = rng.uniform(-lim, lim, shape)
        else:
            d[name] = rng.uniform(-0.1, 0.1, shape)
    return M.pack(d).astype(np.float32)


def test_forward_matches_cpu_network():
    g = CollisionSolver3d()
    w = _weights(0)
    g.setLearnedWeights(w)
    np.testing.assert_array_equal(g.learnedWeights(), w)
    rng = np.random.default_rng(1)
    x = rng.normal(0, 3, (1000, 24)).astype(np.float32)  # 15 full tiles + 40 rows
    out = g.learnedInfer(x)
    ref = M.forward(w, x, emulate_bf16=True)
    scale = np.abs(ref).max()
    err = np.abs(out - ref).max() / scale
    print(f"[mlp] forward max error / max |out| = {err:.2e}")
    assert err < 1e-2
    # a transposed W2 (a layout mistake) in the CPU network is far off
    p = M.unpack(w.astype(np.float64).copy())
    p["W2"] = p["W2"].T.copy()
    assert np.abs(out - M.forward(M.pack(p), x, emulate_bf16=True)).max() / scale > 5e-2
    # a row's outputs do not depend on the other rows of its tile
    np.testing.assert_array_equal(g.learnedInfer(x[:1]), out[:1])
    np.testing.assert_array_equal(g.learnedInfer(x[937:]), out[937:])
    g.close()


def _layer_checks(grad, ref):
    """per-layer cosine similarity and norm ratio of two flat gradients"""
    res = {}
    off = 0
    for name, shape in M.SHAPES:
        n = int(np.prod(shape))
        a, b = grad[off:off + n].astype(np.float64), ref[off:off + n]
        off += n
        na, nb = np.linalg.norm(a), np.linalg.norm(b)
        res[name] = (float(a @ b / (na * nb)) if na > 0 and nb > 0 else 1.0,
                     float(na / nb) if nb > 0 else 1.0)
    return res


def _layers_ok(res):
    return all(c > 0.99 and 0.95 < r < 1.05 for c, r in res.values())


def test_gradient_matches_cpu_network():
    g = CollisionSolver3d()
    w = _weights(2)
    g.setLearnedWeights(w)
    rng = np.random.default_rng(3)
    n = 300  # four full tiles and a ragged one
    x = rng.normal(0, 2, (n, 24)).astype(np.float32)
    e = (rng.random(n) < 0.4).astype(np.float32)
    t = (rng.random(n) * e).astype(np.float32)
    loss, grad = g.learnedGrad(x, e, t)
    # the CPU network sees the same BF16-rounded inputs and weights
    xq = M.bf16(x).astype(np.float64)
    wq = w.astype(np.float64)
    p = M.unpack(wq)
    for k in ("W1", "W2", "W3"):
        p[k][...] = M.bf16(p[k].astype(np.float32))
    lref, gref = M.loss_and_grad(wq, xq, e, t)
    res = _layer_checks(grad, gref)
    print(f"[mlp] loss {loss:.6f} vs {lref:.6f}; per layer (cos, norm ratio) {res}")
    assert abs(loss - lref) <= 1e-2 * abs(lref)
    assert _layers_ok(res)
    for plant in ("no_sigmoid_deriv", "w2_untransposed"):
        _, gp = M.loss_and_grad(wq, xq, e, t, _plant=plant)
        assert not _layers_ok(_layer_checks(grad, gp)), plant
    # the gradient is a descent direction of the device's own loss with the
    # predicted first-order decrease: a step of 2% of the loss along -grad
    gn = float(grad.astype(np.float64) @ grad)
    eps = 0.02 * loss / gn
    g.setLearnedWeights((w - eps * grad).astype(np.float32))
    loss2, _ = g.learnedGrad(x, e, t)
    print(f"[mlp] step along -grad: loss {loss:.6f} -> {loss2:.6f} (first order {loss - eps * gn:.6f})")
    assert 0.5 * eps * gn < loss - loss2 < 1.5 * eps * gn
    g.close()


@pytest.fixture(scope="module")
def trained():
    """a network trained on the plain CCD's own sub-tests of config 1 at 64x64"""
    sc = S.self_collision_cloth(n=64, steps=12)
    g = solver_for_scene(sc)
    g.learnedInit(11)
    w0 = g.learnedWeights()
    g.learnedCollect(capacity=1 << 21, stride=8)
    for _ in range(12):
        g.step()
    x, e, t = g.learnedDataset()
    losses = g.learnedTrain(epochs=20, batch=8192, lr=1e-3, seed=5)
    w = g.learnedWeights()
    # the same initial weights, data and seed give the same weights
    g.setLearnedWeights(w0)
    g.learnedTrain(epochs=20, batch=8192, lr=1e-3, seed=5)
    w_again = g.learnedWeights()
    g.close()
    return dict(scene=sc, x=x, e=e, t=t, losses=losses, w=w, w0=w0, w_again=w_again)


def test_dataset_labels_follow_the_cubic(trained):
    x, e, t = trained["x"], trained["e"], trained["t"]
    sc = trained["scene"]
    n = len(e)
    print(f"[mlp] dataset {n} samples, {int(e.sum())} with a root")
    assert n > 10000 and 0 < e.sum() < n
    assert set(np.unique(e)) <= {0.0, 1.0}
    assert np.all((t >= 0) & (t <= 1)) and np.all(t[e == 0] == 0)
    # rebuild the sub-tests from their features (the cubic only depends on
    # differences) and check the labelled time of impact is a root
    h, dt = sc.thickness, sc.dt
    rng = np.random.default_rng(0)
    ok = tot = 0
    for i in rng.choice(np.flatnonzero(e == 1), size=min(400, int(e.sum())), replace=False):
        xo = x[i, :12].astype(np.float64).reshape(4, 3) * h
        v = x[i, 12:].astype(np.float64).reshape(4, 3) * h / dt
        a, b, c, d = M.cubic_coeffs(xo, v)
        r = np.roots([a, b, c, d])
        r = r[np.abs(r.imag) <= 1e-9 * dt].real
        tot += 1
        ok += bool(np.any(np.abs(r - t[i] * dt) <= 1e-3 * dt))
    print(f"[mlp] labelled roots matching the rebuilt cubic: {ok}/{tot}")
    assert ok >= 0.9 * tot


def test_training_reduces_loss_deterministically(trained):
    losses = trained["losses"]
    print(f"[mlp] epoch losses {np.round(losses, 5).tolist()}")
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.8 * losses[0]
    np.testing.assert_array_equal(trained["w"], trained["w_again"])
    assert not np.array_equal(trained["w"], trained["w0"])


def test_learned_form_is_bitwise_plain_config1(trained):
    """Config 1 at 64x64 with the trained network ordering and seeding the
    CCD: every count, hit pair, zone and state bit equals the CPU oracle's."""
    sc = S.self_collision_cloth(n=64, steps=12)
    st = run_parity(sc, steps=12, verbose=True, gpu_kw=dict(learned=1),
                    gpu_setup=lambda s: s.setLearnedWeights(trained["w"]))
    assert st["hit_pairs"] > 0


def test_learned_and_mixed_form_is_bitwise_plain_config4():
    """Config 4 geometry at 96x96 with an untrained network and the mixed
    precision screens: any weights leave the results unchanged."""
    sc = S.combined_scene(n=96, nstrings=60, steps=60)
    st = run_parity(sc, steps=60, check_every=5, verbose=True,
                    gpu_kw=dict(learned=1, precision=1), gpu_setup=lambda s: s.learnedInit(3))
    assert st["hit_pairs"] > 0


def test_trained_seed_and_statistics(trained):
    """The statistics of the learned form on config 1 (64x64): every scored
    sub-test with a root gets a located Newton root next to the closed form
    (the two agree to well within 1e-4 dt; ill-conditioned cubics separate
    them by about 1e-6 dt), the trained network finds most roots and its
    seeds need fewer Newton iterations than the untrained network's and
    than the bracket midpoint."""
    res = {}
    for tag, w in (("trained", trained["w"]), ("untrained", trained["w0"])):
        g = solver_for_scene(trained["scene"], learned=1)
        g.setLearnedWeights(w)
        for _ in range(12):
            g.step()
        res[tag] = g.learnedStats()
        g.close()
        r = res[tag]
        print(f"[mlp] {tag}: {r}; iterations per root seeded "
              f"{r['newton_iters'] / max(1, r['newton_roots']):.3f} vs midpoint "
              f"{r['newton_iters_mid'] / max(1, r['newton_roots']):.3f}; accuracy "
              f"{r['agree'] / max(1, r['scored']):.4f}")
    tr = res["trained"]
    assert tr["scored"] > 0 and tr["exists"] > 0
    assert tr["newton_roots"] >= 0.99 * tr["exists"]
    assert tr["max_root_err"] < 1e-4
    assert tr["newton_iters"] < res["untrained"]["newton_iters"]
    assert tr["newton_iters"] < tr["newton_iters_mid"]
