"""Pins the absolute GT value of the oracle's pairing (its Miller-loop sign convention for the
negative BLS parameter x) through a reduced Tate pairing computed by an independent Miller loop.

The oracle's pairing is the optimal ate pairing of weighted_btx/src/final_exponentiation.rs:
229-273 (loop over |x|, conjugated because x < 0) followed by the reference's final
exponentiation, whose hard chain raises to 3 Lambda (final_exponentiation.rs:308-336).  The
Hess-Smart-Vercauteren relation for the ate pairing with T = t - 1 = x (Hess, Smart,
Vercauteren, "The Eta Pairing Revisited", IEEE Trans. IT 52(10), 2006, Theorem 2) fixes that
pairing against the reduced Tate pairing t(Q, P) = f_{r,Q}(P)^((p^12 - 1)/r):

    f_{x,Q}(P)^((p^12-1)/r * c)  ==  t(Q, P)^L,   c = sum_{i<12} x^(11-i) p^i,  L = (x^12 - 1)/r,

because f_{x^12,Q} = prod_i f_{x,[x^i]Q}^(x^(11-i)) = f_{x,Q}^c ([x]Q is the p-power Frobenius of
Q in G2, and f_{x,pi(Q)}(P) = f_{x,Q}(P)^p for P over Fp) and f_{x^12,Q} = f_{1+Lr,Q} = f_{r,Q}^L
(normalised Miller functions; [r]Q = O).  So blst's pairing e = f_{x,Q}(P)^(3 (p^12-1)/r) obeys
e^c == t(Q, P)^(3L).  Dropping the final conjugation of the ate loop inverts e, and the check
then fails (t(Q, P)^(6L) != 1 for Q, P of order r).

The Tate Miller loop below runs on the untwisted curve E(Fp12): y^2 = x^3 + 4 in affine
coordinates with exact slopes, vertical lines and normalised functions; it shares only the
field arithmetic with the oracle (pinned on its own in test_oracle.py).  CPU only.
"""
import random

from helpers import O

F6Z = O.F6_ZERO


def fp12_from_fp(a):
    return (((a % O.P, 0), O.F2_ZERO, O.F2_ZERO), F6Z)


def fp12_from_fp2(a):
    return ((a, O.F2_ZERO, O.F2_ZERO), F6Z)


def fp12_sub(a, b):
    return (O.f6_sub(a[0], b[0]), O.f6_sub(a[1], b[1]))


def fp12_neg(a):
    return (O.f6_neg(a[0]), O.f6_neg(a[1]))


W = (F6Z, (O.F2_ONE, O.F2_ZERO, O.F2_ZERO))  # w, with w^2 = v and w^6 = xi
W_INV2 = O.f12_inv(O.f12_mul(W, W))
W_INV3 = O.f12_inv(O.f12_pow(W, 3))
FOUR = fp12_from_fp(4)


def untwist(q_aff):
    """E'(Fp2): y^2 = x^3 + 4 xi  ->  E(Fp12): y^2 = x^3 + 4, (x, y) -> (x / w^2, y / w^3)."""
    x, y = q_aff
    return (O.f12_mul(fp12_from_fp2(x), W_INV2), O.f12_mul(fp12_from_fp2(y), W_INV3))


def on_curve(pt):
    x, y = pt
    return O.f12_mul(y, y) == O.f12_add(O.f12_mul(O.f12_mul(x, x), x), FOUR)


def ec_add(a, b):
    """Affine addition on E(Fp12); None is the point at infinity."""
    if a is None:
        return b
    if b is None:
        return a
    (x1, y1), (x2, y2) = a, b
    if x1 == x2:
        if O.f12_add(y1, y2) == O.f12_add(O.F12_ONE, fp12_neg(O.F12_ONE)):  # y1 == -y2
            return None
        lam = O.f12_mul(O.f12_mul(fp12_from_fp(3), O.f12_mul(x1, x1)), O.f12_inv(O.f12_add(y1, y1)))
    else:
        lam = O.f12_mul(fp12_sub(y2, y1), O.f12_inv(fp12_sub(x2, x1)))
    x3 = fp12_sub(fp12_sub(O.f12_mul(lam, lam), x1), x2)
    return (x3, fp12_sub(O.f12_mul(lam, fp12_sub(x1, x3)), y1))


def ec_mul(a, k):
    r = None
    for bit in bin(k)[2:]:
        r = ec_add(r, r)
        if bit == "1":
            r = ec_add(r, a)
    return r


def tate_miller(q, p, n):
    """Normalised Miller function f_{n,Q}(P) for Q on E(Fp12) and P = (xp, yp) over Fp, by
    double-and-add with lines l_{T,S}(P) = (yp - yT) - lam (xp - xT) and verticals
    v_R(P) = xp - xR; numerator and denominator are kept apart and divided once."""
    xp, yp = fp12_from_fp(p[0]), fp12_from_fp(p[1])
    num, den = O.F12_ONE, O.F12_ONE
    t = q

    def step(t, s):
        nonlocal num, den
        (xt, yt), (xs, ys) = t, s
        if xt == xs and O.f12_add(yt, ys) == fp12_sub(O.F12_ONE, O.F12_ONE):
            num = O.f12_mul(num, fp12_sub(xp, xt))  # vertical line; the sum is O
            return None
        if t == s:
            lam = O.f12_mul(O.f12_mul(fp12_from_fp(3), O.f12_mul(xt, xt)), O.f12_inv(O.f12_add(yt, yt)))
        else:
            lam = O.f12_mul(fp12_sub(ys, yt), O.f12_inv(fp12_sub(xs, xt)))
        num = O.f12_mul(num, fp12_sub(fp12_sub(yp, yt), O.f12_mul(lam, fp12_sub(xp, xt))))
        x3 = fp12_sub(fp12_sub(O.f12_mul(lam, lam), xt), xs)
        y3 = fp12_sub(O.f12_mul(lam, fp12_sub(xt, x3)), yt)
        den = O.f12_mul(den, fp12_sub(xp, x3))
        return (x3, y3)

    for bit in bin(n)[3:]:
        num, den = O.f12_mul(num, num), O.f12_mul(den, den)
        t = step(t, t)
        if bit == "1":
            t = step(t, q)
    assert t is None, "n Q must be the point at infinity"
    return O.f12_mul(num, O.f12_inv(den))


def reduced_tate(q_aff_twist, p_aff):
    q = untwist(q_aff_twist)
    return O.f12_pow(tate_miller(q, p_aff, O.R), (O.P**12 - 1) // O.R)


def hsv_exponents():
    x = O.X
    assert (x**12 - 1) % O.R == 0
    c = sum(x ** (11 - i) * O.P**i for i in range(12))
    L = (x**12 - 1) // O.R
    return c % O.R, L % O.R


def test_untwisted_points_lie_in_the_trace_zero_subgroup():
    """The untwist puts G2 on E(Fp12) with pi(Q) = [p]Q = [x]Q (p = x mod r)."""
    q = untwist(O.G2_GEN)
    assert on_curve(q)
    frob = (O.f12_frobenius(q[0], 1), O.f12_frobenius(q[1], 1))
    xq = ec_mul(q, O.X_ABS)
    assert frob == (xq[0], fp12_neg(xq[1]))  # [x]Q = -[|x|]Q
    assert O.P % O.R == O.X % O.R


def test_hsv_exponents_nondegenerate():
    c, L = hsv_exponents()
    assert c == 12 * pow(O.X, 11, O.R) % O.R  # c = k p^(k-1) = 12 x^11 mod r
    assert c != 0 and L != 0


def test_oracle_pairing_matches_tate_through_hsv():
    """e_oracle(P, Q)^c == t(Q, P)^(3L) for the generators and for random multiples; the
    opposite sign convention (no conjugation) would give e^-1 and fail."""
    c, L = hsv_exponents()
    rnd = random.Random(12)
    a, b = rnd.randrange(1, O.R), rnd.randrange(1, O.R)
    cases = [(O.G1_GEN, O.G2_GEN),
             (O.G1.to_affine(O.G1.mul(O.g1_gen_jac(), a)), O.G2.to_affine(O.G2.mul(O.g2_gen_jac(), b)))]
    for p, q in cases:
        e = O.pairing(p, q)
        t = reduced_tate(q, p)
        assert t != O.F12_ONE and O.f12_pow(t, O.R) == O.F12_ONE
        assert O.f12_pow(e, c) == O.f12_pow(t, 3 * L % O.R)
        # the unconjugated loop's value, e^-1, does not satisfy the relation
        assert O.f12_pow(O.f12_conj(e), c) != O.f12_pow(t, 3 * L % O.R)


def test_sign_check_catches_a_planted_unconjugated_loop(monkeypatch):
    """Plant the mistake in the oracle (drop the final conjugation of the ate loop): the HSV
    relation with the Tate pairing then fails."""
    c, L = hsv_exponents()
    orig = O.miller_loop
    monkeypatch.setattr(O, "miller_loop", lambda pairs: O.f12_conj(orig(pairs)))
    e_wrong = O.pairing(O.G1_GEN, O.G2_GEN)
    t = reduced_tate(O.G2_GEN, O.G1_GEN)
    assert O.f12_pow(e_wrong, c) != O.f12_pow(t, 3 * L % O.R)
