"""The device Barth limiter divides once per sign group of Delta_2 instead of
once per face (barth_psi_d / barth_psi_f in the kernels).  This checks that
grouping against the reference's face-by-face form (include/Numerics.hpp:
179-236) in IEEE double and single arithmetic, on random faces and on the
cases the argument rests on: exact zero numerators, ties, zero and NaN
deltas, deltas far below and near eps, mixed signs."""
import numpy as np


def reference_psi(delta, vmax, vmin, v, eps, dt):
    """limiter = min over faces of min(1, num / Delta_2), sequential std::min."""
    one = dt(1.0)
    lim = dt(np.finfo(dt).max)
    for d in delta:
        s = dt(int(d > 0) - int(d < 0))
        d2 = dt(s * (dt(abs(d)) + eps))
        if d2 > 0:
            q = dt(dt(vmax - v) / d2)
            fc = q if q < one else one
        elif d2 < 0:
            q = dt(dt(vmin - v) / d2)
            fc = q if q < one else one
        else:
            fc = one
        lim = fc if fc < lim else lim
    return lim


def grouped_psi(delta, vmax, vmin, v, eps, dt):
    hi = dt(0.0)
    lo = dt(0.0)
    for d in delta:
        hi = np.fmax(hi, d)
        lo = np.fmin(lo, d)
    one = dt(1.0)
    pp = np.fmin(one, dt(vmax - v) / dt(hi + eps)) if hi > 0 else one
    pm = np.fmin(one, dt(vmin - v) / dt(lo - eps)) if lo < 0 else one
    return dt(np.fmin(pp, pm))


def _cases(dt, rng, n):
    eps = np.finfo(dt).eps
    special = [dt(0.0), dt(-0.0), dt(eps), dt(-eps), dt(eps / 4), dt(-eps / 4), dt(1e-30), dt(-1e-30),
               dt(0.5), dt(-0.5), dt(np.nan), dt(3.0)]
    for i in range(n):
        v = dt(rng.choice([0.0, 1.0, rng.standard_normal()]))
        nb = [dt(rng.choice([v, v + 1, v - 1, rng.standard_normal()])) for _ in range(4)]
        vmax = max([v] + nb)
        vmin = min([v] + nb)
        if i % 3 == 0:
            delta = [dt(rng.choice(special)) for _ in range(4)]
        elif i % 3 == 1:
            d = dt(rng.standard_normal() * 10.0 ** rng.integers(-20, 2))
            delta = [d, d, dt(-d), dt(rng.choice(special))]  # ties
        else:
            delta = [dt(rng.standard_normal() * 10.0 ** rng.integers(-20, 2)) for _ in range(4)]
        yield delta, vmax, vmin, v


def test_grouped_limiter_equals_face_by_face():
    for dt in (np.float64, np.float32):
        rng = np.random.default_rng(7)
        eps = np.finfo(dt).eps
        n = 0
        with np.errstate(all="ignore"):
            for delta, vmax, vmin, v in _cases(dt, rng, 20000):
                a = reference_psi(delta, vmax, vmin, v, eps, dt)
                b = grouped_psi(delta, vmax, vmin, v, eps, dt)
                assert a == b, (dt, delta, vmax, vmin, v, a, b)
                n += 1
        assert n == 20000


def test_grouping_would_catch_a_wrong_group_choice():
    """Taking the smallest |Delta_2| of a group instead of the largest (a
    plausible mistake) gives different limiters on these cases."""
    dt = np.float64
    eps = np.finfo(dt).eps
    delta = [dt(0.5), dt(2.0), dt(-0.1), dt(-3.0)]
    vmax, vmin, v = dt(1.0), dt(-1.0), dt(0.0)
    ref = reference_psi(delta, vmax, vmin, v, eps, dt)
    wrong = min(1.0, (vmax - v) / (0.5 + eps), (vmin - v) / (-0.1 - eps))
    assert ref == grouped_psi(delta, vmax, vmin, v, eps, dt) and ref != wrong
