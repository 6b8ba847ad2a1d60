"""Constants of the FP64 exponential that the straight-line chemistry kernels inline (hip_emit.cpp, exp64).

e^x = 2^n e^r with n = rint(x / ln 2) and r = x - n ln 2 (Cody and Waite's two-constant reduction: n ln2_hi is exact
for |n| < 2^21 because ln2_hi keeps only the leading 32 bits of ln 2), and e^r for |r| <= ln(2)/2 by a degree-10
polynomial: the Chebyshev interpolant of e^r on that interval (computed in 60-digit arithmetic, within a small
factor of the minimax error; its constant coefficient rounds to 1 in FP64, so e^0 = 1 exactly). Its worst relative
error is 4.3e-16 in FP64 Horner form (about two ulps; degree 11: 1.6e-16, one FMA more; degree 9: 1.9e-14 and a
constant coefficient off 1). Prints the constants as C hex literals and the polynomial's worst relative error over a
dense grid, evaluated in FP64 Horner form.

Run: python exp_coefficients.py
"""
import mpmath as mp

mp.mp.dps = 60
DEGREE = 10
a = mp.log(2) / 2


def hexd(x):
    return float(x).hex()


def main():
    ln2 = mp.log(2)
    hi = mp.mpf(float(ln2))
    hi = mp.floor(hi * 2 ** 32) / 2 ** 32  # leading 32 bits
    lo = ln2 - hi
    # Chebyshev interpolation of e^r on [-a, a] at DEGREE+1 Chebyshev nodes, converted to the monomial basis.
    n = DEGREE + 1
    nodes = [a * mp.cos(mp.pi * (2 * k + 1) / (2 * n)) for k in range(n)]
    A = mp.matrix([[x ** j for j in range(n)] for x in nodes])
    c = mp.lu_solve(A, mp.matrix([mp.e ** x for x in nodes]))
    coef = [c[j] for j in range(n)]
    cd = [float(x) for x in coef]
    worst = 0
    for i in range(20001):
        r = -a + 2 * a * i / 20000
        rf = float(r)
        p = cd[-1]
        for j in range(DEGREE - 1, -1, -1):
            p = p * rf + cd[j]  # FP64 Horner (rounding per step, as the kernel's FMAs at worst)
        worst = max(worst, abs((mp.mpf(p) - mp.e ** mp.mpf(rf)) / mp.e ** mp.mpf(rf)))
    print("ln2_hi", hexd(hi), "ln2_lo", hexd(lo), "log2e", hexd(1 / ln2))
    print("coefficients c0..c%d:" % DEGREE, ", ".join(hexd(x) for x in coef))
    print("worst relative error of the FP64 polynomial on [-ln2/2, ln2/2]: %.3g" % worst)


if __name__ == "__main__":
    main()
