"""The two-register QMS layout on the oracle (the commented bm_enes_old code of
qms.cuh:72-85, 134-136, 201-202; the reference has no working version of it,
so it is pinned against the reference-pinned one-register path):
* the reference's own comparison hooks: its one-register metro_step carries
  commented "debug only (comparison with 2enereg)" draws (qms.cuh:134-136,
  433, 479) that make it consume the two-register chain's random stream; with
  them, the one-register chain's return codes, pcg32 state after every step,
  E/X lists and reverse counters equal the two-register chain's, over exact
  ([-1, 3]) and inexact ([0, 1]) energy ranges, restorations that fail (-1),
  r = 2..5 and several betas and seeds (the plain one-register chain draws
  fewer numbers);
* W with E_old read from its register equals, bit for bit, the one-register
  W with the classical E_old set to the register's value (every (E_old,
  E_new, state) case at r = 3);
* the chain reproduces the one-register thermal averages: over 24 seeds of
  400 steps at the c1 settings (beta 1, E in [-1, 3], r = 3), mean energy and
  acceptance agree within 3 combined standard errors;
* layout: 3 + 2r + 1 qubits, E_old below E_new, acceptance on top."""
import numpy as np
import pytest

from oracle import pyoracle as po


def test_layout():
    q = po.Qms(1.0, 3, 5, -1.0, 3.0, 10, 100, False, two_registers=True)
    f = q.fields
    assert f.nqubits == 3 + 2 * 3 + 1 and f.two_registers == 1
    assert list(f.bm_enes_old[:3]) == [3, 4, 5] and list(f.bm_enes[:3]) == [6, 7, 8] and f.bm_acc == 9
    assert f.W_mask_Eold == 0b111 << 3 and f.W_mask_Enew == 0b111 << 6
    q1 = po.Qms(1.0, 3, 5, -1.0, 3.0, 10, 100, False)
    assert q1.fields.nqubits == 7 and q1.fields.two_registers == 0


@pytest.mark.parametrize("beta", [0.3, 2.0])
def test_W_reads_E_old_from_the_register(beta):
    r, rng = 3, np.random.default_rng(11)
    two = po.Qms(beta, r, 5, -1.0, 3.0, 10, 100, False, two_registers=True)
    one = po.Qms(beta, r, 5, -1.0, 3.0, 10, 100, False)
    for e_old in range(1 << r):
        # one-register state: index = syst (3) | E_new (r) << 3 | acc << (3 + r)
        n1 = 3 + r + 1
        re1, im1 = rng.standard_normal(1 << n1), rng.standard_normal(1 << n1)
        # the same amplitudes in the two-register layout with E_old = e_old
        re2, im2 = np.zeros(1 << (n1 + r)), np.zeros(1 << (n1 + r))
        for i in range(1 << n1):
            syst, rest = i & 7, i >> 3
            j = syst | (e_old << 3) | (rest << (3 + r))
            re2[j], im2[j] = re1[i], im1[i]
        one.set_state(re1, im1)
        one.fields.curr_E_old = e_old
        one.apply_W()
        two.set_state(re2, im2)
        two.fields.curr_E_old = (e_old + 3) % (1 << r)   # must not be used
        two.apply_W()
        o_re, o_im = one.state()
        t_re, t_im = two.state()
        for i in range(1 << n1):
            syst, rest = i & 7, i >> 3
            j = syst | (e_old << 3) | (rest << (3 + r))
            assert t_re[j] == o_re[i] and t_im[j] == o_im[i]
        mask = np.ones(1 << (n1 + r), bool)
        mask[[(i & 7) | (e_old << 3) | ((i >> 3) << (3 + r)) for i in range(1 << n1)]] = False
        assert not np.any(t_re[mask]) and not np.any(t_im[mask])


def test_thermal_averages_match_one_register():
    stats = {}
    for two in (False, True):
        means, accs = [], []
        for seed in range(1, 25):
            q = po.Qms(1.0, 3, 4000 + seed, -1.0, 3.0, 10, 100, False, two_registers=two)
            acc, _ = q.run(400, 1, 50)
            E, _ = q.measures()
            means.append(np.mean(E))
            accs.append(acc / 400)
        stats[two] = [(np.mean(v), np.std(v, ddof=1) / np.sqrt(len(v))) for v in (means, accs)]
    for k in range(2):
        (m1, s1), (m2, s2) = stats[False][k], stats[True][k]
        assert abs(m1 - m2) < 3 * np.hypot(s1, s2), (k, m1, s1, m2, s2)


CMP_CASES = [  # beta, r, E range, max_reverse, steps, seeds
    (1.0, 3, (-1.0, 3.0), 100, 300, (1, 2, 3, 4)),
    (1.0, 4, (-1.0, 3.0), 100, 150, (42, 7)),
    (3.0, 3, (0.0, 1.0), 2, 300, (11, 12)),      # failed restorations
    (0.5, 4, (-1.0, 3.0), 2, 200, (13,)),
    (2.0, 2, (0.0, 1.0), 3, 300, (21, 22)),
    (1.0, 5, (-1.0, 3.0), 100, 40, (5,)),
]


@pytest.mark.parametrize("case", CMP_CASES, ids=[f"b{c[0]}-r{c[1]}-E{c[2][0]}_{c[2][1]}-rev{c[3]}" for c in CMP_CASES])
def test_two_registers_equal_comparison_draw_chain(case):
    beta, r, (emin, emax), max_rev, steps, seeds = case
    fails = 0
    for seed in seeds:
        two = po.Qms(beta, r, seed, emin, emax, 10, max_rev, True, two_registers=True)
        cmp_ = po.Qms(beta, r, seed, emin, emax, 10, max_rev, True, compare_draws=True)
        plain = po.Qms(beta, r, seed, emin, emax, 10, max_rev, True)
        lib = po.lib()
        for q in (two, cmp_, plain):
            lib.so_qms_init_state(q.p)
        s0 = 0
        plain.run(steps, 1, 20)
        rets = []
        for s in range(steps):
            take = (s > s0 + 20) and ((s - s0) % 1 == 0)
            a = lib.so_qms_metro_step(two.p, 1 if take else 0)
            b = lib.so_qms_metro_step(cmp_.p, 1 if take else 0)
            assert a == b, (seed, s)
            assert two.rng_state() == cmp_.rng_state(), (seed, s)
            if a < 0:
                for q in (two, cmp_):
                    lib.so_qms_init_state(q.p)
                s0 = s + 1
            rets.append(a)
        assert two.measures() == cmp_.measures()
        assert two.reverse_counters() == cmp_.reverse_counters()
        assert two.rng_state() != plain.rng_state()   # the plain chain draws fewer numbers
        fails += rets.count(-1)
    if max_rev <= 3 and r == 3:
        assert fails > 0
