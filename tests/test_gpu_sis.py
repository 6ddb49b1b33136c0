"""GPU parity of sequential indicator simulation against the CPU oracle.

With FP64 solves the weights are the reference's bit for bit, so every
realisation must equal the oracle's cell for cell (categories and masks).
In the mixed mode a weight differs from FP64 by the refinement residual;
a category flips only when u * sum lands within that difference of a
threshold, and a flip travels along the path to the cells that use it, so
the test asks for at least 99% agreement.  Correlogram systems always take
the FP64 kernels and must match exactly in both modes.
"""
import numpy as np
import pytest

import oracle_binding as O
from cases import Cov, Grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def geo():
    import hpgl_amd.geo_bsd as g

    return g


@pytest.fixture(autouse=True)
def _mixed(geo):
    geo.set_precision("mixed")
    yield
    geo.set_precision("mixed")


def _hard(grid, K, frac, seed):
    rng = np.random.default_rng(seed)
    n = grid.size
    idx = rng.choice(n, max(1, int(frac * n)), replace=False)
    mask = np.zeros(n, np.uint8)
    mask[idx] = 1
    cat = np.zeros(n, np.uint8)
    cat[idx] = rng.integers(0, K, idx.size)
    return np.asfortranarray(cat.reshape(grid.shape, order="F")), np.asfortranarray(mask.reshape(grid.shape, order="F"))


def _data(geo, K, max_n=12, radii=(6, 6, 3)):
    out = []
    for k in range(K):
        cm = geo.CovarianceModel(type=1 if k % 2 == 0 else 0, ranges=(8.0 + 2 * k, 6.0, 4.0),
                                 angles=(15.0 * k, 0.0, 0.0), sill=1.0, nugget=0.05 * (k % 2))
        out.append({"cov_model": cm, "radiuses": radii, "max_neighbours": max_n + 2 * k})
    return out


def _marg(K):
    return {2: [0.35, 0.65], 3: [0.2, 0.3, 0.5], 4: [0.1, 0.2, 0.3, 0.4]}[K]


def _run_both(geo, g, K, seed, marg, frac=0.03, mask=None, correlogram=True, data=None):
    cat, hm = _hard(g, K, frac, seed)
    data = data if data is not None else _data(geo, K)
    print("sis case K=%d seed=%d grid=%s lvm=%s" % (K, seed, g.shape, not np.isscalar(marg[0])), flush=True)
    gp = geo.sis_simulation(geo.IndProperty(cat, hm, K), g, [dict(d) for d in data], seed, marg,
                            use_correlogram=correlogram, mask=mask)
    od, om = O.sis_simulation((cat, hm, K), g, data, seed, marg, use_correlogram=correlogram, mask=mask)
    return gp, od, om, cat, hm


@pytest.mark.parametrize("K,seed", [(3, 12345), (2, 3241347), (4, 777)])
def test_sis_fp64_equals_oracle(geo, K, seed):
    geo.set_precision("fp64")
    g = Grid(40, 32, 12)
    gp, od, om, cat, hm = _run_both(geo, g, K, seed, _marg(K))
    assert np.array_equal(gp.mask, om)
    assert np.array_equal(gp.data, od), np.argwhere(gp.data != od)[:10]
    assert np.array_equal(gp.data[hm == 1], cat[hm == 1])


@pytest.mark.parametrize("K", [2, 3])
def test_sis_mixed_agrees_with_oracle(geo, K):
    # 20 neighbours and more: the orders at which mixed mode factors in FP32
    g = Grid(48, 40, 12)
    gp, od, om, cat, hm = _run_both(geo, g, K, 99 + K, _marg(K), data=_data(geo, K, max_n=20))
    assert np.array_equal(gp.mask, om)
    agree = float(np.mean(gp.data == od))
    assert agree >= 0.99, agree
    st = geo.last_stats()
    assert st["solves_fp32"] > 0


@pytest.mark.parametrize("correlogram", [False, True])
def test_sis_lvm_equals_oracle(geo, correlogram):
    geo.set_precision("fp64")
    g = Grid(36, 30, 10)
    K = 3
    rng = np.random.default_rng(11)
    raw = rng.uniform(0.05, 1.0, (K,) + g.shape)
    lvm = [np.asfortranarray((raw[k] / raw.sum(axis=0)).astype(np.float32)) for k in range(K)]
    lvm[0][0, 0, 0] = 0.0
    lvm[2][1, 0, 0] = 1.0
    gp, od, om, cat, hm = _run_both(geo, g, K, 2024, lvm, correlogram=correlogram)
    assert np.array_equal(gp.mask, om)
    assert np.array_equal(gp.data, od), np.argwhere(gp.data != od)[:10]


def test_sis_correlogram_is_fp64_in_mixed_mode(geo):
    g = Grid(30, 30, 8)
    K = 2
    rng = np.random.default_rng(5)
    m0 = np.asfortranarray(rng.uniform(0.1, 0.9, g.shape).astype(np.float32))
    lvm = [m0, np.asfortranarray((1 - m0).astype(np.float32))]
    gp, od, om, cat, hm = _run_both(geo, g, K, 31, lvm, correlogram=True)
    assert np.array_equal(gp.data, od)
    assert geo.last_stats()["solves_fp32"] == 0


def test_sis_simulation_mask_and_no_harddata(geo):
    geo.set_precision("fp64")
    g = Grid(32, 28, 8)
    K = 3
    sim = np.zeros(g.shape, np.uint8, order="F")
    sim[:, ::3, :] = 1
    gp, od, om, cat, hm = _run_both(geo, g, K, 8080, _marg(K), mask=sim)
    assert np.array_equal(gp.mask, om)
    assert np.array_equal(gp.data[om == 1], od[om == 1])
    # without hard data the conditioning set starts empty
    data = _data(geo, K)
    e = geo.sis_simulation(geo.IndProperty(cat, hm, K), g, [dict(d) for d in data], 8080, _marg(K),
                           use_harddata=False)
    z = np.zeros_like(cat)
    eo, eom = O.sis_simulation((z, z, K), g, data, 8080, _marg(K))
    assert np.array_equal(e.data, eo)
    assert np.array_equal(e.mask, eom)


def test_sis_batch_equals_single_calls(geo):
    g = Grid(24, 20, 6)
    K = 3
    cat, hm = _hard(g, K, 0.05, 1)
    data = _data(geo, K, max_n=8, radii=(4, 4, 2))
    seeds = [11, 12345, 999999]
    batch = geo.sis_simulation_batch(geo.IndProperty(cat, hm, K), g, [dict(d) for d in data], seeds, _marg(K))
    for s, b in zip(seeds, batch):
        one = geo.sis_simulation(geo.IndProperty(cat, hm, K), g, [dict(d) for d in data], s, _marg(K))
        assert np.array_equal(b.data, one.data)
        assert np.array_equal(b.mask, one.mask)


def test_sis_rejects_one_category(geo):
    g = Grid(8, 8, 2)
    cat, hm = _hard(g, 1, 0.1, 3)
    with pytest.raises(Exception, match="two categories"):
        geo.sis_simulation(geo.IndProperty(cat, hm, 1), g, _data(geo, 1), 5, [1.0])
