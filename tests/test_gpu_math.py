"""The engine's own FP64 device routines, evaluated through nb_eval_math.

The near-singular and singular passes use their own sin/cos (Cody-Waite
reduction with pi/2 in three parts, Taylor series on |r| <= pi/4), natural
logarithm (m 2^e, atanh series), 1/sqrt and 1/x (hardware estimate and two
Newton steps) instead of the library's.  They are checked here against
numpy's extended precision (x86 80-bit long double; its sin, cos and log
are accurate to about 1e-19) over the argument ranges the passes use:
|x| <= 1e4 for sin/cos (angles and k r), [1e-12, 1e12] for the logarithm,
[1e-20, 1e8] for 1/sqrt (squared distances) and [1e-12, 1e12] for 1/x.
Bars: sin/cos 3e-16 absolute, log 4e-16 relative, 1/sqrt and 1/x 3e-16
relative (about one unit in the last place); special values as documented.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "GPU test without a GPU"
    from nihu_amd import nbem
    return nbem


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).cuda()


def _ld(x):
    return np.asarray(x, dtype=np.longdouble)


def test_sincos64(nb):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.uniform(-np.pi, np.pi, 200000),
        rng.uniform(-1e4, 1e4, 200000),
        np.exp(rng.uniform(np.log(1e-12), np.log(1e4), 100000)) * rng.choice([-1.0, 1.0], 100000),
        # around the quadrant boundaries n pi/4, where the reduction switches
        ((np.arange(-4000, 4001) * (np.pi / 4))[:, None] + np.array([-1e-9, 0.0, 1e-9])[None, :]).ravel(),
        [0.0, -0.0, 1e-300, np.pi / 2, np.pi, 1e4],
    ])
    y = nb.eval_math("sincos", _dev(x)).cpu().numpy()
    es = np.abs(_ld(y[:, 0]) - np.sin(_ld(x))).max()
    ec = np.abs(_ld(y[:, 1]) - np.cos(_ld(x))).max()
    print("sincos64 over %d arguments: max abs error sin %.2e cos %.2e" % (len(x), es, ec), flush=True)
    assert es <= 3e-16 and ec <= 3e-16
    # odd / even symmetry is exact, and sin(0) = 0 exactly
    ys = nb.eval_math("sincos", _dev(-x)).cpu().numpy()
    assert np.array_equal(ys[:, 0], -y[:, 0]) and np.array_equal(ys[:, 1], y[:, 1])
    s0 = nb.eval_math("sincos", _dev([0.0, np.nan, np.inf])).cpu().numpy()
    assert s0[0, 0] == 0.0 and s0[0, 1] == 1.0
    assert np.isnan(s0[1]).all() and np.isnan(s0[2]).all()


def test_log64(nb):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        np.exp(rng.uniform(np.log(1e-12), np.log(1e12), 400000)),
        1.0 + rng.uniform(-1e-3, 1e-3, 50000),
        np.sqrt(0.5) * (1.0 + rng.uniform(-1e-6, 1e-6, 50000)),   # the mantissa switch
        np.ldexp(1.0, np.arange(-40, 41)).astype(np.float64),
    ])
    y = nb.eval_math("log", _dev(x)).cpu().numpy()
    ref = np.log(_ld(x))
    err = np.abs(_ld(y) - ref) / np.maximum(np.abs(ref), np.finfo(np.float64).tiny)
    # log near 1 is small: there the absolute error is the meaningful one
    near1 = np.abs(x - 1.0) < 1e-2
    print("log64: max rel error %.2e (away from 1), max abs error near 1 %.2e"
          % (err[~near1].max(), np.abs(_ld(y[near1]) - ref[near1]).max()), flush=True)
    assert err[~near1].max() <= 4e-16
    assert np.abs(_ld(y[near1]) - ref[near1]).max() <= 4e-18 + 4e-16 * np.abs(ref[near1]).max()
    assert nb.eval_math("log", _dev([1.0])).item() == 0.0
    sp = nb.eval_math("log", _dev([0.0, -1.0, np.inf, np.nan])).cpu().numpy()
    assert sp[0] == -np.inf and np.isnan(sp[1]) and sp[2] == np.inf and np.isnan(sp[3])


@pytest.mark.parametrize("fn,lo,hi", [("rsqrt", 1e-20, 1e8), ("rcp", 1e-12, 1e12)])
def test_rsqrt64_rcp64(nb, fn, lo, hi):
    rng = np.random.default_rng(2)
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), 500000))
    if fn == "rcp":
        x = x * rng.choice([-1.0, 1.0], len(x))
    y = nb.eval_math(fn, _dev(x)).cpu().numpy()
    ref = 1.0 / np.sqrt(_ld(x)) if fn == "rsqrt" else 1.0 / _ld(x)
    err = (np.abs(_ld(y) - ref) / np.abs(ref)).max()
    print("%s64: max rel error %.2e" % (fn, err), flush=True)
    assert err <= 3e-16
    # powers of four (rsqrt) and of two (rcp) come out exactly
    p = np.ldexp(1.0, np.arange(-40, 41, 2)).astype(np.float64)
    ye = nb.eval_math(fn, _dev(p)).cpu().numpy()
    assert np.array_equal(ye, 1.0 / np.sqrt(p) if fn == "rsqrt" else 1.0 / p)
