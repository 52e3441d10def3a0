"""Restatement of the reference's Remez algorithm (AlgRemez,
lib/Tools/Remez/alg_remez.cpp) and of RationalApprox
(lib/Tools/RationalApprox/rationalapprox.cpp) in mpmath.

TEST INFRASTRUCTURE ONLY.  Importable from tests/ (the checker of the
library's own quad-precision Remez); the product never imports it.

The reference computes in MPFR (bigfloat, alg_remez.hpp / bigfloat.hpp) at
`precision` decimal digits; GMP/MPFR are not in this image, so the
reference's translation unit cannot be compiled here and this restatement
uses mpmath at the same number of digits.  Pinned in
tests/test_oracle_remez.py by the Chebyshev alternation theorem (which fixes
the minimax approximation uniquely), the reference test's own check
(tests/test_RationalApprox.cpp: r(0.5) against pow(0.5, a/b), and applying
x^(1/2) twice), Zolotarev's closed form for x^(-1/2), and the PFE identity.
"""
import math

import mpmath as mp


class AlgRemez:
    """alg_remez.cpp:22-44 (constructor) and the methods cited below."""

    JMAX = 10000

    def __init__(self, lower, upper, precision):
        self.prec = int(precision)
        mp.mp.dps = self.prec
        self.apstrt = mp.mpf(lower)
        self.apend = mp.mpf(upper)
        self.apwidt = self.apend - self.apstrt
        self.tolerance = mp.mpf("1e-15")
        self.n = self.d = 0
        self.found_roots = False

    # ---------------------------------------------------------------- function
    def func(self, x):
        """alg_remez.cpp:516-551: x^(power_num/power_den)."""
        if self.power_num == 0:
            return mp.mpf(1)
        if x == 1:
            return mp.mpf(1)
        return mp.power(x, mp.mpf(self.power_num) / mp.mpf(self.power_den))

    def approx(self, x):
        """alg_remez.cpp:469-481: P(x)/Q(x), Q monic."""
        n, d, p = self.n, self.d, self.param
        yn = p[n]
        for i in range(n - 1, -1, -1):
            yn = x * yn + p[i]
        yd = x + p[n + d]
        for i in range(n + d - 1, n, -1):
            yd = x * yd + p[i]
        return yn / yd

    def get_err(self, x):
        """alg_remez.cpp:485-499: relative error and its sign."""
        f = self.func(x)
        e = self.approx(x) - f
        if f != 0:
            e /= f
        if e < 0:
            return -e, -1
        return e, 1

    # ---------------------------------------------------------------- driver
    def generate_approx(self, num_degree, den_degree, power_num, power_den):
        """alg_remez.cpp:117-203."""
        self.n, self.d = num_degree, den_degree
        self.neq = self.n + self.d + 1
        self.power_num, self.power_den = power_num, power_den
        self.param = [mp.mpf(0)] * (self.neq)
        self.xx = [mp.mpf(0)] * (self.neq + 2)
        self.mm = [mp.mpf(0)] * (self.neq + 1)
        self.spread = mp.mpf("1e37")
        self.iter = 0
        self.initial_guess()
        step = self.stpini()
        while self.spread > self.tolerance:
            self.iter += 1
            self.equations()
            if self.delta < self.tolerance:
                raise RuntimeError("Delta too small, try increasing precision")
            self.search(step)
        err, _ = self.get_err(self.mm[0])
        if not self.root():
            raise RuntimeError("root finding failed")
        self.found_roots = True
        return err

    def initial_guess(self):
        """alg_remez.cpp:296-320 (ncheb = neq)."""
        # the reference forms r in double precision (M_PI, cos, exp)
        ncheb = self.neq
        a = float(ncheb)
        self.mm[0] = self.apstrt
        for i in range(1, ncheb):
            r = 0.5 * (1 - math.cos((math.pi * i) / a))
            r = (math.exp(r) - 1.0) / (math.exp(1.0) - 1.0)
            self.mm[i] = self.apstrt + mp.mpf(r) * self.apwidt
        self.mm[ncheb] = self.apend
        a = 2.0 * ncheb
        for i in range(ncheb + 1):
            r = 0.5 * (1 - math.cos(math.pi * (2 * i + 1) / a))
            r = (math.exp(r) - 1.0) / (math.exp(1.0) - 1.0)
            self.xx[i] = self.apstrt + mp.mpf(r) * self.apwidt

    def stpini(self):
        """alg_remez.cpp:324-334."""
        neq = self.neq
        step = [mp.mpf(0)] * (neq + 1)
        self.xx[neq + 1] = self.apend
        self.delta = mp.mpf(0.25)
        step[0] = self.xx[0] - self.apstrt
        for i in range(1, neq):
            step[i] = self.xx[i] - self.xx[i - 1]
        step[neq] = step[neq - 1]
        return step

    def search(self, step):
        """alg_remez.cpp:338-432: march to the error extremum between
        adjacent zeros, then move the zeros to level the extrema."""
        neq = self.neq
        meq = neq + 1
        yy = [mp.mpf(0)] * meq
        eclose = mp.mpf("1e30")
        farther = mp.mpf(0)
        xx0 = self.apstrt
        for i in range(meq):
            steps = 0
            xx1 = self.xx[i]
            if i == meq - 1:
                xx1 = self.apend
            xm = self.mm[i]
            ym, emsign = self.get_err(xm)
            q = step[i]
            xn = xm + q
            if xn < xx0 or xn >= xx1:
                q = -q
                xn, yn, ensign = xm, ym, emsign
            else:
                yn, ensign = self.get_err(xn)
                if yn < ym:
                    q = -q
                    xn, yn, ensign = xm, ym, emsign
            while yn >= ym:
                steps += 1
                if steps > 10:
                    break
                ym, xm, emsign = yn, xn, ensign
                a = xm + q
                if a == xm or a <= xx0 or a >= xx1:
                    break
                xn = a
                yn, ensign = self.get_err(xn)
            self.mm[i] = xm
            yy[i] = ym
            if eclose > ym:
                eclose = ym
            if farther < ym:
                farther = ym
            xx0 = xx1
        q = farther - eclose
        if eclose != 0:
            q /= eclose
        if q >= self.spread:
            self.delta *= 0.5
        self.spread = q
        for i in range(neq):
            q = yy[i + 1]
            if q != 0:
                q = yy[i] / q - 1
            else:
                q = mp.mpf(0.0625)
            if q > 0.25:
                q = mp.mpf(0.25)
            q *= self.mm[i + 1] - self.mm[i]
            step[i] = q * self.delta
        step[neq] = step[neq - 1]
        for i in range(neq):
            xm = self.xx[i] - step[i]
            if xm <= self.apstrt or xm >= self.apend:
                continue
            if xm <= self.mm[i]:
                xm = mp.mpf(0.5) * (self.mm[i] + self.xx[i])
            if xm >= self.mm[i + 1]:
                xm = mp.mpf(0.5) * (self.mm[i + 1] + self.xx[i])
            self.xx[i] = xm

    def equations(self):
        """alg_remez.cpp:436-465: P(x_i) - f(x_i) Q(x_i) = 0 at the neq
        current zeros, Q monic; solved by Gaussian elimination (simq,
        :555-683; mpmath's LU with partial pivoting here)."""
        n, d, neq = self.n, self.d, self.neq
        A = mp.matrix(neq, neq)
        B = mp.matrix(neq, 1)
        for i in range(neq):
            x = self.xx[i]
            y = self.func(x)
            z = mp.mpf(1)
            col = 0
            for _ in range(n + 1):
                A[i, col] = z
                z *= x
                col += 1
            z = mp.mpf(1)
            for _ in range(d):
                A[i, col] = -y * z
                z *= x
                col += 1
            B[i] = y * z
        X = mp.lu_solve(A, B)
        self.param = [X[i] for i in range(neq)]

    # ---------------------------------------------------------------- roots
    @staticmethod
    def poly_eval(x, poly, size):
        """alg_remez.cpp:726-731."""
        f = poly[size]
        for i in range(size - 1, -1, -1):
            f = f * x + poly[i]
        return f

    @staticmethod
    def poly_diff(x, poly, size):
        """alg_remez.cpp:734-739."""
        df = size * poly[size]
        for i in range(size - 1, 0, -1):
            df = df * x + i * poly[i]
        return df

    def rtnewt(self, poly, i, x1, x2, xacc):
        """alg_remez.cpp:743-758: Newton from the midpoint of [x1, x2]."""
        rtn = mp.mpf(0.5) * (x1 + x2)
        for _ in range(self.JMAX):
            f = self.poly_eval(rtn, poly, i)
            df = self.poly_diff(rtn, poly, i)
            dx = f / df
            rtn -= dx
            if abs(dx) < xacc:
                return rtn
        return mp.mpf(0)

    def root(self):
        """alg_remez.cpp:686-722: roots of P and Q by Newton from -49999.5
        with synthetic division (deflation) after each root."""
        n, d = self.n, self.d
        upper, lower, tol = mp.mpf(1), mp.mpf(-100000), mp.mpf("1e-20")
        poly = [mp.mpf(0)] * (self.neq + 1)
        for i in range(n + 1):
            poly[i] = self.param[i]
        self.roots = [mp.mpf(0)] * n
        for i in range(n - 1, -1, -1):
            self.roots[i] = self.rtnewt(poly, i + 1, lower, upper, tol)
            if self.roots[i] == 0:
                return False
            poly[0] = -poly[0] / self.roots[i]
            for j in range(1, i + 1):
                poly[j] = (poly[j - 1] - poly[j]) / self.roots[i]
        poly[d] = mp.mpf(1)
        for i in range(d):
            poly[i] = self.param[n + 1 + i]
        self.poles = [mp.mpf(0)] * d
        for i in range(d - 1, -1, -1):
            self.poles[i] = self.rtnewt(poly, i + 1, lower, upper, tol)
            if self.poles[i] == 0:
                return False
            poly[0] = -poly[0] / self.poles[i]
            for j in range(1, i + 1):
                poly[j] = (poly[j - 1] - poly[j]) / self.poles[i]
        self.norm = self.param[n]
        return True

    # ---------------------------------------------------------------- PFE
    def pfe(self, res, poles, norm):
        """alg_remez.cpp:763-836: residues of norm * prod(x - res)/prod(x -
        poles) in the form norm + sum r_i/(x + p_i); poles returned as
        positive shifts, sorted ascending."""
        n = self.n
        numerator = [mp.mpf(0)] * n
        denominator = [mp.mpf(0)] * n
        numerator[0] = mp.mpf(1)
        denominator[0] = mp.mpf(1)
        for j in range(n):
            for i in range(n - 1, -1, -1):
                numerator[i] *= -res[j]
                denominator[i] *= -poles[j]
                if i > 0:
                    numerator[i] += numerator[i - 1]
                    denominator[i] += denominator[i - 1]
        for i in range(n):
            numerator[i] -= denominator[i]
        out = [mp.mpf(0)] * n
        for i in range(n):
            r = mp.mpf(0)
            for j in range(n - 1, -1, -1):
                r = poles[i] * r + numerator[j]
            for j in range(n - 1, -1, -1):
                if i != j:
                    r /= poles[i] - poles[j]
            out[i] = r * norm
        p = [-x for x in poles]
        order = sorted(range(n), key=lambda i: p[i])
        return [out[i] for i in order], [p[i] for i in order]

    def get_pfe(self):
        """alg_remez.cpp:206-242."""
        r, p = self.pfe(list(self.roots), list(self.poles), self.norm)
        return r, p, self.norm

    def get_ipfe(self):
        """alg_remez.cpp:246-290: the inverse, x^(-pnum/pden)."""
        r, p = self.pfe(list(self.poles), list(self.roots), 1 / self.norm)
        return r, p, 1 / self.norm


class RationalApprox:
    """RationalApprox (rationalapprox.cpp:11-80, rationalapprox.hpp:56-96):
    x^(a/b) and x^(-a/b) on [low, high] as const + sum res/(x + pole)."""

    def __init__(self, num_deg, den_deg, exp_num, exp_den, precision, low, high):
        assert num_deg == den_deg
        self.params = dict(num_deg=num_deg, den_deg=den_deg, exp_num=exp_num,
                           exp_den=exp_den, precision=precision, low=low, high=high)
        rz = AlgRemez(low, high, precision)
        self.error = rz.generate_approx(num_deg, den_deg, exp_num, exp_den)
        r, p, c = rz.get_pfe()
        ir, ip, ic = rz.get_ipfe()
        self.remez = rz
        self.res, self.pole, self.a0 = r, p, c
        self.ires, self.ipole, self.ia0 = ir, ip, ic

    def exponent(self):
        return self.params["exp_num"] / self.params["exp_den"]

    def Residuals(self):
        return [float(x) for x in self.res]

    def Poles(self):
        return [float(x) for x in self.pole]

    def Const(self):
        return float(self.a0)

    def InvResiduals(self):
        return [float(x) for x in self.ires]

    def InvPoles(self):
        return [float(x) for x in self.ipole]

    def InvConst(self):
        return float(self.ia0)

    def rescale(self, min_in, max_in):
        """rationalapprox.cpp:82-103 (coefficients assumed computed on
        [a, 1]; the range is widened to [0.95 min, 1.05 max])."""
        mx = 1.05 * max_in
        e = self.params["exp_num"] / self.params["exp_den"]
        factor = mx ** e
        self.a0 = mp.mpf(float(self.a0) * factor)
        self.res = [mp.mpf(float(r) * (factor * mx)) for r in self.res]
        self.pole = [mp.mpf(float(p) * mx) for p in self.pole]
        factor = mx ** (-e)
        self.ia0 = mp.mpf(float(self.ia0) * factor)
        self.ires = [mp.mpf(float(r) * (factor * mx)) for r in self.ires]
        self.ipole = [mp.mpf(float(p) * mx) for p in self.ipole]
