"""The block-form exchange ratio the estimators use (blk_ratio in
rbm_pp_amd/csrc/rbmpp_meas_blk_impl.h, big_ratio in rbmpp_meas_blk_big.hip),
checked on the CPU against the determinant ratio it stands for.

With the normal projection every X is [[0, F], [-F^T, 0]], F(u, d) =
Sl(site_u, site_d) for up electrons u and down electrons d (Sl(a, b) the
Slater_Elm entry of row a, column b + N), so the Pfaffian ratio of an
exchange (pfupdate_two.f90:27-88) is det F' / det F.  The kernels form it
from B = F^-1, Yh = Gdd B, Xh(d, u) = B[d] . Guu[., u] and
W(d, u) = Yh[d] . Guu[., u] plus three Slater scalars; this test builds F'
explicitly and compares."""
import numpy as np
import pytest


def blk_ratio(yh, xh, w, b, delta, eps, gdd):
    """The kernels' formula, term for term."""
    a1 = yh + delta * b
    a2 = xh - eps * b
    yc = w - gdd - delta + delta * xh - eps * a1
    return a1 * a2 - yc * b


@pytest.mark.parametrize("N,seed", [(16, 1), (36, 2), (100, 3), (144, 4)])
def test_block_exchange_ratio_matches_determinants(N, seed):
    rng = np.random.default_rng(seed)
    np_ = N // 2
    Sl = rng.standard_normal((N, N))
    sites = rng.permutation(N)
    su_list, sd_list = sites[:np_], sites[np_:]  # electron u sits at su_list[u], d at sd_list[d]
    F = Sl[np.ix_(su_list, sd_list)]
    B = np.linalg.inv(F)  # B[d, u]
    Gdd = Sl[np.ix_(sd_list, sd_list)]  # Gdd(d, d') = Sl(s_d, s_d')
    Guu = Sl[np.ix_(su_list, su_list)]  # Guu(u', u) = Sl(s_u', s_u)
    Yh = Gdd @ B
    detF = np.linalg.det(F)
    worst = 0.0
    for _ in range(40):
        u, d = rng.integers(np_), rng.integers(np_)
        su, sd = su_list[u], sd_list[d]
        # the exchange: up electron u moves to sd, down electron d to su
        su2, sd2 = su_list.copy(), sd_list.copy()
        su2[u], sd2[d] = sd, su
        F2 = Sl[np.ix_(su2, sd2)]
        want = np.linalg.det(F2) / detF
        xh = B[d] @ Guu[:, u]
        w = Yh[d] @ Guu[:, u]
        delta = Sl[sd, su] - Sl[sd, sd]
        eps = Sl[su, su] - Sl[su, sd]
        got = blk_ratio(Yh[d, u], xh, w, B[d, u], delta, eps, Gdd[d, d])
        scale = max(1.0, abs(want))
        worst = max(worst, abs(got - want) / scale)
    # both sides carry the conditioning of F: random Gaussian F of this size
    # keeps the difference near 1e-12
    assert worst < 1e-9, worst


def test_block_exchange_ratio_is_one_for_the_identity_move():
    """Exchanging an electron with itself is not a move; the closest check:
    when su and sd carry the same Slater rows and columns, F' = F."""
    rng = np.random.default_rng(7)
    N, np_ = 20, 10
    Sl = rng.standard_normal((N, N))
    sites = rng.permutation(N)
    su_list, sd_list = sites[:np_], sites[np_:]
    u, d = 3, 5
    su, sd = su_list[u], sd_list[d]
    # make sites su and sd indistinguishable to Sl (equal rows and columns)
    Sl[sd, :] = Sl[su, :]
    Sl[:, sd] = Sl[:, su]
    F = Sl[np.ix_(su_list, sd_list)]
    B = np.linalg.inv(F)
    Gdd = Sl[np.ix_(sd_list, sd_list)]
    Guu = Sl[np.ix_(su_list, su_list)]
    Yh = Gdd @ B
    got = blk_ratio(Yh[d, u], B[d] @ Guu[:, u], Yh[d] @ Guu[:, u], B[d, u], Sl[sd, su] - Sl[sd, sd],
                    Sl[su, su] - Sl[su, sd], Gdd[d, d])
    assert got == pytest.approx(1.0, abs=1e-9)


def test_sampler_rank2_update_matches_inverse():
    """The sampler's accepted move in block form (rbmpp_sampler.hip, the
    pending update applied by blk_upd1 / blk_upd4): F' = F + e_ua r^T + c e_db^T
    with c(ua) = 0, y = r^T B, x = B c (old B), R = (1 + y_ua)(1 + B[db].c)
    - (y.c) B[db][ua] and
      B'[d][u] = B[d][u] + B[d][ua] al_u + x_d be_u,
      al = -((1 + B[db].c) y - (y.c) B[db]) / R,
      be = -((1 + y_ua) B[db] - B[db][ua] y) / R,
    against det F'/det F and inv(F') on random matrices."""
    rng = np.random.default_rng(11)
    for np_ in (8, 25, 50, 128):
        F = rng.standard_normal((np_, np_)) + 3.0 * np.eye(np_)
        B = np.linalg.inv(F)  # B[d, u]
        for _ in range(10):
            ua, db = rng.integers(np_), rng.integers(np_)
            r = rng.standard_normal(np_)  # change of row ua (over the down electrons d)
            c = rng.standard_normal(np_)  # change of column db (over the up electrons u)
            c[ua] = 0.0
            F2 = F + np.outer(np.eye(np_)[ua], r) + np.outer(c, np.eye(np_)[db])
            y = r @ B          # y_u = sum_d r_d B[d, u]
            brow = B[db]       # B[db][u]
            gam, bet = y[ua], brow[ua]
            dlt, eps_ = brow @ c, y @ c
            R = (1.0 + gam) * (1.0 + dlt) - eps_ * bet
            assert R == pytest.approx(np.linalg.det(F2) / np.linalg.det(F), rel=1e-9)
            al = -((1.0 + dlt) * y - eps_ * brow) / R
            be = -((1.0 + gam) * brow - bet * y) / R
            x = B @ c          # x_d = sum_u B[d, u] c_u
            B2 = B + np.outer(B[:, ua], al) + np.outer(x, be)
            want = np.linalg.inv(F2)
            assert np.abs(B2 - want).max() <= 1e-9 * np.abs(want).max()
