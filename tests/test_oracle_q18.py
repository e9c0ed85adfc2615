"""Pins the oracle's reference-faithful pre pass (orc_ray_messages_ex with
brick entries), written out by hand.

The reference's pre pass (rayStartComputingMessagesFromBrickPre,
gvdbkernels.cu:689-819) is called once per active brick and restores its
visibility carry only when the accumulated tau is non-zero (:708-715).  Its
FP32 arithmetic flushes denormals (--use_fast_math, CMakeLists.txt:192), so
a tau is exactly 0 when one of its products, a = 0.398942 e^{-u^2/2},
psi = a / (3 sigma), b = V (1 - n) or tau = b psi, is below 2^-126.  Until
the first tau that survives, every brick starts again from V = 1; T and the
infinity term V_n (:1196-1208) use these brick-local visibilities, while the
post pass carries its visibility along the whole ray (:844-856) and takes
R_i = T - sum_{j<=i} tau_j.
"""
import math

import numpy as np
import pytest

import oracle

F_EPS = 1e-4
FTZ = 2.0 ** -126
SIG = 0.05  # sigma (m), 3 sigma = 0.15 m


def _ftz_parts(u, V, n):
    a = 0.398942 * math.exp(-0.5 * u * u)
    ps = a / (3 * SIG)
    b = V * (1 - n)
    tau = b * ps
    return tau, min(a, ps, b, tau) >= FTZ


def _psi(s, d):
    u = abs(d - s) / (3 * SIG)
    return 0.398942 * math.exp(-0.5 * u * u) / (3 * SIG)


def _post(n, l, s, d, T):
    """post pass with full-ray visibility and a given T: p_i"""
    L, P, out = 0.0, 0.0, []
    taus = []
    for i in range(len(n)):
        L += l[i] * math.log(n[i])
        V = math.exp(L)
        tau = V * (1 - n[i]) * _psi(s[i], d)
        taus.append(tau)
        R = T - (P + tau)
        mu1 = P + V * _psi(s[i], d)
        mu0 = P + R / n[i]
        out.append(min(max(mu1 / (mu1 + mu0 + 1e-7), F_EPS), 1 - F_EPS))
        P += tau
    return np.array(out), sum(taus)


def test_two_bricks_visibility_restarts_before_the_surface():
    """Brick 1 (records 0-2) lies 14+ likelihood widths in front of the
    surface, so its taus flush to 0; brick 2 (records 3-5) holds the surface.
    The pre pass enters brick 2 with V = 1 again: T = sum_{i=3..5}
    (prod_{3<=j<=i} n_j^l_j) (1 - n_i) psi_i.  Catches: visibility carried into
    brick 2 (T smaller by prod_{j<3} n_j^l_j)."""
    d = 2.0
    s = [d - 2.4, d - 2.25, d - 2.1, d - 0.1, d, d + 0.05]  # u = 16, 15, 14, 0.67, 0, 0.33
    n = [0.9, 0.8, 0.7, 0.6, 0.5, 0.4]
    l = [1.0, 0.5, 1.2, 1.0, 0.7, 1.0]
    bstart = [1, 0, 0, 1, 0, 0]
    V3 = n[3] ** l[3]
    V4 = V3 * n[4] ** l[4]
    V5 = V4 * n[5] ** l[5]
    T = V3 * (1 - n[3]) * _psi(s[3], d) + V4 * (1 - n[4]) * _psi(s[4], d) + V5 * (1 - n[5]) * _psi(s[5], d)
    for i in range(3):
        assert not _ftz_parts(abs(d - s[i]) / (3 * SIG), 1.0, n[i])[1]
    p, Tg = oracle.ray_messages(n, l, s, d, [SIG, 0, 0], bstart=bstart, ran_out=False)
    assert Tg == pytest.approx(T, rel=1e-12)
    want, Tcarry = _post(n, l, s, d, T)
    np.testing.assert_allclose(p, want, rtol=1e-10)
    # without brick entries the pre pass carries: T is the post pass's own sum
    pc, Tc = oracle.ray_messages(n, l, s, d, [SIG, 0, 0], ran_out=False)
    assert Tc == pytest.approx(Tcarry, rel=1e-12)
    assert Tg / Tc == pytest.approx(1.0 / (n[0] * n[1] ** 0.5 * n[2] ** 1.2), rel=1e-9)
    assert np.abs(p - pc).max() > 1e-2


def test_carry_from_the_first_surviving_tau():
    """The first tau that survives lies in brick 1 (record 2): from the entry
    of brick 2 on, the pre pass carries brick 1's visibility sum (which
    began at brick 1's entry).  Catches: a restart at every brick whatever
    the accumulated tau."""
    d = 2.0
    s = [d - 2.4, d - 2.25, d - 1.5, d - 0.1, d, d + 0.05]  # record 2: u = 10
    n = [0.9, 0.8, 0.7, 0.6, 0.5, 0.4]
    l = [1.0, 1.0, 1.0, 1.0, 1.0, 1.0]
    bstart = [1, 0, 0, 1, 0, 0]
    L = np.cumsum(np.log(n))  # brick 1 began at record 0
    taus = [math.exp(L[i]) * (1 - n[i]) * _psi(s[i], d) for i in range(6)]
    assert not _ftz_parts(16.0, math.exp(L[0]), n[0])[1] and not _ftz_parts(15.0, math.exp(L[1]), n[1])[1]
    assert _ftz_parts(10.0, math.exp(L[2]), n[2])[1]
    T = sum(taus[2:])
    p, Tg = oracle.ray_messages(n, l, s, d, [SIG, 0, 0], bstart=bstart)
    assert Tg == pytest.approx(T, rel=1e-12)
    np.testing.assert_allclose(p, _post(n, l, s, d, T)[0], rtol=1e-10)


def test_infinity_term_uses_the_last_bricks_visibility():
    """The walk ran out with d beyond the last voxel (gvdbkernels.cu:1196-1208)
    and no tau survives before the last brick: T += V_n with V_n the
    visibility since the last brick's entry."""
    d = 5.0
    s = [1.0, 1.1, 1.2, 4.5, 4.6, 4.7]  # last brick 2 .. 3.3 likelihood widths in front of d
    n = [0.9, 0.8, 0.7, 0.6, 0.5, 0.4]
    l = [1.0, 1.0, 1.0, 1.0, 1.0, 1.0]
    bstart = [1, 0, 0, 1, 0, 0]
    Vl = np.cumprod(n[3:])
    T = sum(Vl[i] * (1 - n[3 + i]) * _psi(s[3 + i], d) for i in range(3)) + Vl[-1]
    p, Tg = oracle.ray_messages(n, l, s, d, [SIG, 0, 0], bstart=bstart, ran_out=True)
    assert Tg == pytest.approx(T, rel=1e-12)
    np.testing.assert_allclose(p, _post(n, l, s, d, T)[0], rtol=1e-10)


def test_gap_record_starts_a_brick():
    """A gap record (inactive bricks) is followed by a brick entry: the
    visibility sum restarts there, and the gap itself attenuates nothing."""
    d = 2.0
    s = [d - 2.4, d - 2.25, d - 1.0, d - 0.1, d]
    n = [0.9, 0.8, 1.0, 0.6, 0.5]
    l = [1.0, 1.0, 30.0, 1.0, 1.0]
    live = [1, 1, 0, 1, 1]
    bstart = [1, 0, 0, 1, 0]
    V3 = n[3]
    V4 = V3 * n[4]
    T = V3 * (1 - n[3]) * _psi(s[3], d) + V4 * (1 - n[4]) * _psi(s[4], d)
    p, Tg = oracle.ray_messages(n, l, s, d, [SIG, 0, 0], live=live, bstart=bstart)
    assert Tg == pytest.approx(T, rel=1e-12)


@pytest.mark.parametrize("b", [0.5, 0.1, 1e-3])
def test_flush_to_zero_threshold_in_closed_form(b):
    """One record (V = 1: l = 0): its tau enters T exactly when the smallest
    of a = 0.398942 e^{-u^2/2}, psi = a / (3 sigma) and tau = b psi is at
    least 2^-126, i.e. when 0.398942 m e^{-u^2/2} >= 2^-126 with
    m = min(1, 1 / (3 sigma), b / (3 sigma)):
    u*^2 = 2 (126 ln 2 + ln(0.398942 m)).  Records at u* - 0.002 and
    u* + 0.002 fall on either side."""
    n = 1.0 - b  # b = V (1 - n)
    m = min(1.0, 1.0 / (3 * SIG), b / (3 * SIG))
    ustar = math.sqrt(2 * (126 * math.log(2) + math.log(0.398942 * m)))
    for du, inside in ((-0.002, True), (0.002, False)):
        u = ustar + du
        d = 1.0 + u * 3 * SIG
        _, Tg = oracle.ray_messages([n], [0.0], [1.0], d, [SIG, 0, 0], bstart=[1])
        want = b * _psi(1.0, d)
        smallest = 0.398942 * m * math.exp(-0.5 * u * u)
        if inside:
            assert Tg == pytest.approx(want, rel=1e-9) and smallest >= FTZ
        else:
            assert Tg == 0.0 and smallest < FTZ
