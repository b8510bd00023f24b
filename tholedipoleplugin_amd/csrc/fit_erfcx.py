"""Rational approximation of the scaled complementary error function erfcx(x) = e^{x^2} erfc(x) on
[0, 6], used by the FP64 real-space radial functions of the Ewald option (ew_erfc in tdp_ewald.hip).

Prints the coefficients of P(x) / Q(x), P and Q of degree 9 with P(0) = Q(0) = 1, found by the
Sanathanan-Koerner iteration (linear least squares of P - f Q = 0 weighted by 1 / (f Q_previous),
which converges to a near-equal-ripple relative error) over 400 Chebyshev points, in 50-digit
arithmetic (mpmath).  The fit's relative error is 2e-18; evaluated in FP64 (Horner, one division)
it stays below 1e-15, which tests/test_erfcx_fit.py checks against SciPy's erfc.

Run: python tholedipoleplugin_amd/csrc/fit_erfcx.py
"""
import mpmath as mp

X, DEG, NPT, ITERS = 6.0, 9, 400, 12


def fit():
    mp.mp.dps = 50
    f = lambda x: mp.exp(x * x) * mp.erfc(x)
    xs = [mp.mpf(X) * (1 - mp.cos(mp.pi * (k + mp.mpf(1) / 2) / NPT)) / 2 for k in range(NPT)]
    fs = [f(x) for x in xs]
    qprev = [mp.mpf(1)] * NPT
    for _ in range(ITERS):
        # unknowns p_1..p_n, q_1..q_n with p_0 = q_0 = 1 (erfcx(0) = 1)
        A = mp.matrix(NPT, 2 * DEG)
        b = mp.matrix(NPT, 1)
        for k, (x, fx) in enumerate(zip(xs, fs)):
            s = 1 / (fx * qprev[k])
            for j in range(1, DEG + 1):
                A[k, j - 1] = s * x ** j
                A[k, DEG + j - 1] = -s * fx * x ** j
            b[k] = s * (fx - 1)
        sol = mp.qr_solve(A, b)[0]
        p = [mp.mpf(1)] + [sol[j] for j in range(DEG)]
        q = [mp.mpf(1)] + [sol[DEG + j] for j in range(DEG)]
        qprev = [mp.polyval(q[::-1], x) for x in xs]
    err = max(abs(mp.polyval(p[::-1], x) / mp.polyval(q[::-1], x) / fx - 1) for x, fx in zip(xs, fs))
    return [float(a) for a in p], [float(a) for a in q], float(err)


if __name__ == "__main__":
    p, q, err = fit()
    print("// relative error of the fit: %.2e" % err)
    print("P = {%s}" % ", ".join("%.17g" % a for a in p))
    print("Q = {%s}" % ", ".join("%.17g" % a for a in q))
