"""The differentiation workload's parameter fit (`bench.fit_along_line`) on synthetic losses whose gradients
misjudge the response, as a gradient with the tree walk's decisions held fixed does over a long horizon."""
import numpy as np

import bench

P0 = np.array([45.0, 1.2, 0.11])
SCALE = np.array([45.0, 1.2, 0.11])


def _loss(p):
    # smooth, nonlinear in all three parameters, increasing along its gradient
    a = np.array([0.8, 1.5, 9.0])
    return 2.0e19 * float(np.exp(a @ ((p - P0) / SCALE)))


def _grad(p, h=1e-7):
    g = np.zeros(3)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h * max(abs(p[k]), 1.0)
        g[k] = (_loss(p + e) - _loss(p - e)) / (2 * e[k])
    return g



def test_exact_gradient_of_a_linear_loss_lands_in_one_step():
    ev = lambda p: (float(3.0 * p[0] + 2.0 * p[1] - p[2]), np.array([3.0, 2.0, -1.0]))
    h = bench.fit_along_line(ev, P0, SCALE, 200.0, 2)
    assert abs(h[1]["loss_kg"] - 200.0) < 1e-9
    assert abs(h[-1]["loss_kg"] - 200.0) < 1e-9


def test_misjudged_gradient_converges_on_the_secant():
    target = 2.0 * _loss(P0)
    for factor in (0.35, 0.67, 1.5, 3.0):  # the gradient's size off by these factors, its direction exact
        ev = lambda p, f=factor: (_loss(p), f * _grad(p))
        h = bench.fit_along_line(ev, P0, SCALE, target, 4)
        assert abs(h[-1]["loss_kg"] / target - 1.0) < 1e-3, (factor, [x["loss_kg"] / target for x in h])


def test_gauss_newton_alone_would_miss_where_the_line_search_does_not():
    # the previous fit (three Gauss-Newton steps on the misjudged gradient) for comparison
    target = 2.0 * _loss(P0)
    grad = lambda p: 1.5 * _grad(p)
    p = P0.copy()
    for _ in range(3):
        sg = SCALE * grad(p)
        p = p + SCALE * sg * (target - _loss(p)) / float(sg @ sg)
    gn = _loss(p) / target
    h = bench.fit_along_line(lambda q: (_loss(q), grad(q)), P0, SCALE, target, 3)
    assert abs(h[-1]["loss_kg"] / target - 1.0) < abs(gn - 1.0)


def test_a_response_of_the_wrong_sign_is_followed():
    # along the step the loss falls where the gradient says it rises: the secant turns around
    ev = lambda p: (float(-p[0]), np.array([1.0, 0.0, 0.0]))
    h = bench.fit_along_line(ev, P0, SCALE, -40.0, 3)
    assert abs(h[-1]["loss_kg"] + 40.0) < 1e-9


def test_the_previous_fit_history_is_bracketed_and_closed():
    # the shape of the measured fit (start 0.50 of the target, the first step at 1.25): a concave response
    # with a gradient 1.5x too small converges once the bracket exists
    target = 2.0
    ev = lambda p: (1.0 + 1.5 * float(np.tanh((p[0] - 45.0) / 10.0)) * 1.1, np.array([0.1, 0.0, 0.0]))
    h = bench.fit_along_line(ev, P0, SCALE, target, 5)
    assert abs(h[-1]["loss_kg"] / target - 1.0) < 1e-4


def test_zero_gradient_returns_the_start_only():
    h = bench.fit_along_line(lambda p: (1.0, np.zeros(3)), P0, SCALE, 2.0, 3)
    assert len(h) == 1 and h[0]["loss_kg"] == 1.0


def test_a_loss_that_is_not_finite_ends_the_fit():
    calls = []

    def ev(p):
        calls.append(p)
        return (float("nan") if len(calls) > 1 else 1.0), np.array([1.0, 0.0, 0.0])

    h = bench.fit_along_line(ev, P0, SCALE, 2.0, 4)
    assert len(h) == 2 and len(calls) == 2
