"""Third-body efficiency lists longer than any shipped mechanism holds: the
test mechanism written by conftest.extend_efficiencies (gri30 with 9-22
efficiencies per third-body reaction) checked through the oracle against the
closed form of [M] (CPU; the GPU side is tests/test_gpu_long_efficiencies.py).

For a three-body reaction the forward rate of progress is k_f [M] prod C, with
[M] = ctot (d + sum_k (e_k - d) X_k) (ThirdBodyCalc.h:75-91), so extending the
list scales ropf by exactly the ratio of the two [M] sums; every other
reaction's rates stay the same.
"""
import numpy as np

from conftest import extend_efficiencies, mech_file, parse_third_bodies


def test_extended_lists_scale_three_body_rates_by_m_ratio(oracle_mod, tmp_path):
    src = mech_file("gri30")
    dst = str(tmp_path / "gri30_long_eff.mech")
    lists = extend_efficiencies(src, dst)
    lens = sorted(len(e) for _, e in lists.values())
    assert lens[0] >= 9 and lens[-1] <= 22 and len(lists) == len(parse_third_bodies(src))
    assert [e for _, e in lists.values()] == [e for _, e in parse_third_bodies(dst)]
    om0, om1 = oracle_mod.Mech(src), oracle_mod.Mech(dst)
    orig = {}
    with open(src) as f:
        for line in f:
            t = line.split()
            if t and t[0] == "R":
                p = 15
                nr = int(t[p]); p += 1 + 3 * nr
                np_ = int(t[p]); p += 1 + 2 * np_
                orig[int(t[1])] = (int(t[2]), float(t[p]),
                                   [(int(t[p + 2 + 2 * i]), float(t[p + 3 + 2 * i])) for i in range(int(t[p + 1]))])
    g = np.random.default_rng(5)
    K = om0.K
    checked = 0
    for s in range(20):
        T = g.uniform(800.0, 2500.0)
        P = 101325.0 * np.exp(g.uniform(np.log(0.1), np.log(50.0)))
        X = 10.0 ** g.uniform(-8, 0, K)
        X /= X.sum()
        Y = X * om0.mw / (X * om0.mw).sum()
        r0, r1 = om0.rates(T, P, Y), om1.rates(T, P, Y)
        for j, (typ, d, eff0) in orig.items():
            if j not in lists:
                assert r1["ropf"][j] == r0["ropf"][j] and r1["ropr"][j] == r0["ropr"][j]
                continue
            if typ != 1:
                continue
            m0 = d + sum(e * X[k] for k, e in eff0)
            m1 = d + sum(e * X[k] for k, e in lists[j][1])
            assert abs(r1["ropf"][j] - r0["ropf"][j] * m1 / m0) <= 1e-12 * abs(r0["ropf"][j] * m1 / m0)
            checked += 1
    assert checked > 0
