"""Parallel-tempering bookkeeping (OpenAcc/HPT_utilities.c) on the CPU.

The oracle (oracle/hpt.c) is pinned by
  * init_k: the defect box enumerated independently from coordinates, the
    closed-form count of c_r links, and a hand-worked defect_info;
  * the label manager: hand-worked swap sequences (all accepted / none
    accepted, even/odd and serial chains), the order in which uniforms are
    consumed, and permutation invariants over random runs.
The library's host label manager (ostp_pt_*) is then compared with the oracle
bit for bit (exact ints) on random action and random-number sequences; it
makes no GPU call.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

import openstaple_amd as ost
from oracle import Oracle, PT_ACTION, PT_UNIFORM, manage_replica_swaps, metro_swap


# ------------------------------------------------------------------ init_k
def defect_box(shape, map4, def_axis, def_vec):
    """Independent enumeration: logical coordinates of the defect sites."""
    gn = [shape[map4[p]] for p in range(4)]
    ext = [0] * 4
    ext[def_axis] = 1
    perp = [q for q in range(4) if q != def_axis]
    for i, q in enumerate(perp):
        ext[q] = def_vec[i]
    sites = []
    for d in itertools.product(*(range(n) for n in shape)):
        phys = [d[map4[p]] for p in range(4)]
        ok = True
        for p_ in range(4):
            lo = gn[p_] - 1 if p_ == def_axis else 0
            ok = ok and lo <= phys[p_] < lo + ext[p_]
        if ok:
            sites.append(d)
    return sites


CASES = [
    ((4, 4, 4, 4), (0, 1, 2, 3), 3, (2, 3, 4)),
    ((4, 6, 8, 10), (0, 1, 2, 3), 0, (6, 8, 10)),
    ((4, 6, 8, 10), (3, 0, 1, 2), 1, (2, 5, 3)),
    ((6, 4, 4, 2), (1, 2, 3, 0), 2, (3, 4, 1)),
    ((4, 4, 4, 4), (0, 1, 2, 3), 2, (0, 4, 4)),  # empty defect
]


@pytest.mark.parametrize("shape,map4,def_axis,def_vec", CASES)
def test_init_k_planes_match_enumerated_defect(shape, map4, def_axis, def_vec):
    o = Oracle(*shape, par_temp=True)
    conf = o.gauge()
    conf[:] = 7  # K must be overwritten everywhere
    c_r = 0.3125
    o.init_k(conf, c_r, def_axis, def_vec, map4)
    K = o.k_planes(conf)
    dam = map4[def_axis]
    sites = defect_box(shape, map4, def_axis, def_vec)
    expect = np.ones((8, o.sizeh))
    for d in sites:
        idxh = (d[0] + shape[0] * (d[1] + shape[1] * (d[2] + shape[2] * d[3]))) // 2
        expect[2 * dam + sum(d) % 2, idxh] = c_r
    np.testing.assert_array_equal(K, expect)
    # closed form: one c_r link per defect site, |defect| = prod(def_vec)
    assert int((K == c_r).sum()) == math.prod(def_vec)
    # rows of the links are untouched
    per = 152 * o.sizeh
    for lnk in range(8):
        assert np.all(conf[lnk * per:lnk * per + 144 * o.sizeh] == 7)


def test_init_k_defect_info_hand_worked():
    # 4^4, defect on the t boundary, x<2, y<3, z<4: logical d = physical
    o = Oracle(4, 4, 4, 4, par_temp=True)
    conf = o.gauge()
    di = o.init_k(conf, 0.5, 3, (2, 3, 4))
    mn = di[:16].reshape(4, 4)
    mx = di[16:32].reshape(4, 4)
    mnT = di[32:64].reshape(2, 4, 4)
    mxT = di[64:96].reshape(2, 4, 4)
    # sites span d0 in [0,1], d1 in [0,2], d2 in [0,3], d3 = 3; max is exclusive (+1)
    base_min = np.array([0, 0, 0, 3])
    base_max = np.array([2, 3, 4, 4])
    for nu in range(4):
        e = base_min.copy()
        e[nu] -= 1                       # plaquette reach x+nu
        np.testing.assert_array_equal(mn[nu], e)
        np.testing.assert_array_equal(mx[nu], base_max)
        e0 = base_min.copy()
        e0[nu] -= 2                      # 1x2 rectangle reach x+2nu
        np.testing.assert_array_equal(mnT[0, nu], e0)
        e1 = base_min.copy()
        e1[nu] -= 1
        e1[3] -= 1                       # 2x1 rectangle: x+nu and x+mu (mu = t)
        np.testing.assert_array_equal(mnT[1, nu], e1)
        np.testing.assert_array_equal(mxT[0, nu], base_max)
        np.testing.assert_array_equal(mxT[1, nu], base_max)
    assert di[96] == 3 and list(di[97:100]) == [0, 1, 2]


def test_init_k_defect_info_empty_defect_is_zero():
    o = Oracle(4, 4, 4, 4, par_temp=True)
    di = o.init_k(o.gauge(), 0.5, 2, (0, 4, 4))
    assert np.all(di[:96] == 0) and di[96] == 2


def test_init_k_config1_leaves_defect_info_alone():
    o = Oracle(4, 4, 4, 4, par_temp=True)
    conf = o.gauge()
    out = np.full(100, -9, dtype=np.int32)
    from oracle import lib, p
    lib.or_init_k(o.gp, (C.c_int * 4)(0, 1, 2, 3), p(conf), 0.25, 3, (C.c_int * 3)(2, 2, 2), p(out), 1)
    assert np.all(out == -9)
    assert int((o.k_planes(conf) == 0.25).sum()) == 8


# ----------------------------------------------------------- label manager
class Draws:
    """Uniforms from a fixed list; records how many were consumed."""

    def __init__(self, values):
        self.values = list(values)
        self.used = 0

    def __call__(self, _user):
        v = self.values[self.used]
        self.used += 1
        return v


def make_action(n, fn):
    def f(cr_p, label_p, S_p, _user):
        for r in range(n):
            S_p[label_p[r]] = fn(r, cr_p[r])

    return PT_ACTION(f)


def oracle_run(n, evenodd, cr, label, action_fn, draws):
    counters = (C.c_int(0), np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32))
    act = make_action(n, action_fn)
    uni = PT_UNIFORM(draws)
    manage_replica_swaps(n, evenodd, cr, label, act, uni, counters)
    return counters[0].value, counters[1], counters[2]


def test_metropolis_rule():
    assert metro_swap(-1e-300, 0.999) == 1
    assert metro_swap(0.0, 0.999) == 1          # exp(0) = 1 > u
    assert metro_swap(1.0, math.exp(-1.0) - 1e-12) == 1
    assert metro_swap(1.0, math.exp(-1.0) + 1e-12) == 0


def test_evenodd_all_accepted_hand_worked():
    n = 6
    cr = np.linspace(0, 1, n)
    label = np.arange(n, dtype=np.int32)
    # first draw 0.7 >= 0.5 -> swap_order 0: label pairs (0,1)(2,3)(4,5), then
    # the second stage flips the order: (1,2)(3,4).  S_next > S_prev makes
    # Delta S < 0, so every pair is accepted without a further draw.
    draws = Draws([0.7])
    counters = (C.c_int(0), np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32))
    manage_replica_swaps(n, True, cr, label, alternating_action(n, 10.0, 20.0), PT_UNIFORM(draws), counters)
    assert draws.used == 1
    # [0..5] -> (0,1)(2,3)(4,5) -> [1,0,3,2,5,4] -> labels 1<->2, 3<->4 -> [2,0,4,1,5,3]
    assert list(label) == [2, 0, 4, 1, 5, 3]
    assert counters[0].value == 2
    assert list(counters[1]) == [1, 1, 1, 1, 1, 0]
    assert list(counters[2]) == [1, 1, 1, 1, 1, 0]


def alternating_action(n, prev, nxt):
    trial = {"flag": False}

    def f(cr_p, label_p, S_p, _user):
        for r in range(n):
            S_p[label_p[r]] = nxt if trial["flag"] else prev
        trial["flag"] = not trial["flag"]

    return PT_ACTION(f)


def test_evenodd_none_accepted_consumes_one_uniform_per_pair():
    n = 7
    label = np.arange(n, dtype=np.int32)
    # S_next < S_prev: Delta S = +2*(prev-next) = 20 > 0, u = 0.99 > exp(-20): rejected
    draws = Draws([0.2] + [0.99] * 6)  # 0.2 < 0.5 -> swap_order 1: pairs (1,2)(3,4)(5,6), then (0,1)(2,3)(4,5)
    counters = (C.c_int(0), np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32))
    manage_replica_swaps(n, True, np.linspace(0, 1, n), label, alternating_action(n, 10.0, 0.0),
                         PT_UNIFORM(draws), counters)
    assert draws.used == 1 + 3 + 3
    assert list(label) == list(range(n))
    assert counters[0].value == 2
    assert list(counters[1]) == [1, 1, 1, 1, 1, 1, 0]
    assert list(counters[2]) == [0] * n


def test_serial_chain_all_accepted_hand_worked():
    n = 4
    # first draw 0.1 < 0.5 -> swap_order 1: label pairs (0,1), (1,2), (2,3)
    label = np.arange(n, dtype=np.int32)
    draws = Draws([0.1])
    counters = (C.c_int(0), np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32))
    manage_replica_swaps(n, False, np.linspace(0, 1, n), label, alternating_action(n, 0.0, 5.0),
                         PT_UNIFORM(draws), counters)
    # [0,1,2,3] -> 0<->1 [1,0,2,3] -> 1<->2 [2,0,1,3] -> 2<->3 [3,0,1,2]
    assert list(label) == [3, 0, 1, 2]
    assert counters[0].value == 3 and draws.used == 1
    assert list(counters[1]) == [1, 1, 1, 0] and list(counters[2]) == [1, 1, 1, 0]  # on the first label
    # first draw 0.9 -> swap_order 0: (3,2), (2,1), (1,0); counters on the second label
    label = np.arange(n, dtype=np.int32)
    draws = Draws([0.9])
    counters = (C.c_int(0), np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32))
    manage_replica_swaps(n, False, np.linspace(0, 1, n), label, alternating_action(n, 0.0, 5.0),
                         PT_UNIFORM(draws), counters)
    # [0,1,2,3] -> 3<->2 [0,1,3,2] -> 2<->1 [0,2,3,1] -> 1<->0 [1,2,3,0]
    assert list(label) == [1, 2, 3, 0]
    assert list(counters[1]) == [1, 1, 1, 0]


def random_setup(seed, n):
    rng = np.random.default_rng(seed)
    cr = np.sort(rng.random(n))
    coef = rng.normal(size=(n, 3))
    uniforms = rng.random(100000)

    def action(r, c):
        return float(coef[r, 0] * c + coef[r, 1] * c * c + coef[r, 2])

    return cr, action, uniforms


@pytest.mark.parametrize("n,evenodd", [(2, True), (3, True), (6, True), (7, False), (6, False), (2, False)])
def test_oracle_labels_stay_a_permutation(n, evenodd):
    cr, action, uniforms = random_setup(n * 10 + evenodd, n)
    label = np.arange(n, dtype=np.int32)
    draws = Draws(uniforms)
    counters = (C.c_int(0), np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32))
    act = make_action(n, action)
    uni = PT_UNIFORM(draws)
    for _ in range(50):
        manage_replica_swaps(n, evenodd, cr, label, act, uni, counters)
        assert sorted(label) == list(range(n))
    assert np.all(counters[2] <= counters[1])
    assert counters[0].value == 50 * (2 if evenodd else n - 1)


@pytest.mark.parametrize("n,evenodd", [(2, True), (3, True), (6, True), (7, True), (7, False), (6, False), (2, False)])
def test_library_label_manager_matches_oracle(n, evenodd):
    cr, action, uniforms = random_setup(1000 + n * 10 + evenodd, n)
    # oracle
    lab_o = np.arange(n, dtype=np.int32)
    d_o = Draws(uniforms)
    cnt_o = (C.c_int(0), np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32))
    act_o, uni_o = make_action(n, action), PT_UNIFORM(d_o)
    # library
    cr_c = (C.c_double * n)(*cr)
    lab_c = (C.c_int * n)(*range(n))
    rep = ost.RepInfo(n, int(evenodd), 3, (C.c_int * 3)(1, 1, 1), cr_c, lab_c)
    d_l = Draws(uniforms)
    swap_num = C.c_int(0)
    all_l = (C.c_int * n)()
    acc_l = (C.c_int * n)()

    def lib_action(cr_p, label_p, S_p, _user):
        for r in range(n):
            S_p[label_p[r]] = action(r, cr_p[r])

    act_l, uni_l = ost.PT_ACTION_FN(lib_action), ost.PT_UNIFORM_FN(d_l)
    for sweep in range(60):
        manage_replica_swaps(n, evenodd, cr, lab_o, act_o, uni_o, cnt_o)
        ost.lib.ostp_pt_manage_replica_swaps(C.byref(rep), act_l, uni_l, None, C.byref(swap_num), all_l, acc_l)
        assert list(lab_c) == list(lab_o), sweep
        assert d_l.used == d_o.used
    assert swap_num.value == cnt_o[0].value
    assert list(all_l) == list(cnt_o[1]) and list(acc_l) == list(cnt_o[2])
    assert 0 < sum(acc_l) < sum(all_l)  # the random actions exercise both outcomes


def test_library_pt_helpers():
    n = 5
    lab = (C.c_int * n)(4, 2, 0, 1, 3)
    rep = ost.RepInfo(n, 1, 0, (C.c_int * 3)(), (C.c_double * n)(), lab)
    assert ost.lib.ostp_pt_index_of_label0(C.byref(rep)) == 2
    ost.lib.ostp_pt_init_labels(C.byref(rep))
    assert list(lab) == list(range(n))
    assert [ost.lib.ostp_pt_twin_label(lb, 0, n) for lb in range(n)] == [1, 0, 3, 2, -1]
    assert [ost.lib.ostp_pt_twin_label(lb, 1, n) for lb in range(n)] == [-1, 2, 1, 4, 3]
    for ds, u in [(-0.5, 0.99), (0.0, 0.5), (2.0, 0.1), (2.0, 0.2)]:
        assert ost.lib.ostp_pt_metropolis(ds, u) == metro_swap(ds, u)
