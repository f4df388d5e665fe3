"""Exact solution of the 1D Riemann problem of an ideal gas (TEST INFRASTRUCTURE ONLY).

Written from Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics" (3rd ed., 2009), chapter 4: the pressure
functions f_K(p) (eqs. 4.6, 4.7) and their derivatives (4.37), Newton iteration for the star pressure (sec. 4.3.2), the
star velocity (4.9), the sampling of the wave pattern (sec. 4.5, figs. 4.14-4.15) and, for data that generate vacuum
or hold it on one side, the two rarefaction fans around empty space (sec. 4.6), with the conservative cell average over
midpoint sub-samples that luna-cpp's tests use (exact_riemann.hpp:87-101).
"""
import math


class Solver:
    def __init__(self, left, right, gamma):
        self.L, self.R, self.g = tuple(left), tuple(right), gamma
        g = gamma
        rl, ul, pl = self.L
        rr, ur, pr = self.R
        self.cl = math.sqrt(g * pl / rl) if rl > 0 else 0.0
        self.cr = math.sqrt(g * pr / rr) if rr > 0 else 0.0
        self.vacuum = 2 * (self.cl + self.cr) / (g - 1) <= ur - ul or rl == 0 or rr == 0
        if self.vacuum:
            self.p, self.u = 0.0, 0.0
            return
        p = max(1e-12, 0.5 * (pl + pr))
        for _ in range(100):
            fl, dl = self._f(p, self.L, self.cl)
            fr, dr = self._f(p, self.R, self.cr)
            nxt = max(1e-14, p - (fl + fr + ur - ul) / (dl + dr))
            done = 2 * abs(nxt - p) / (nxt + p) < 1e-15
            p = nxt
            if done:
                break
        self.p = p
        fl, _ = self._f(p, self.L, self.cl)
        fr, _ = self._f(p, self.R, self.cr)
        self.u = 0.5 * (ul + ur) + 0.5 * (fr - fl)

    def _f(self, p, K, c):
        g = self.g
        rk, _, pk = K
        if p > pk:  # shock
            A = 2 / ((g + 1) * rk)
            B = (g - 1) / (g + 1) * pk
            q = math.sqrt(A / (p + B))
            return (p - pk) * q, q * (1 - 0.5 * (p - pk) / (B + p))
        # rarefaction
        return (2 * c / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1),
                1 / (rk * c) * (p / pk) ** (-(g + 1) / (2 * g)))

    def _vacuum_sample(self, s):
        """Toro sec. 4.6: the left state, its fan to vacuum (head u_L - c_L, tail u_L + 2 c_L / (g - 1)), empty space,
        the right fan (tail u_R - 2 c_R / (g - 1), head u_R + c_R), the right state; a side with no gas has no fan."""
        g = self.g
        rl, ul, pl = self.L
        rr, ur, pr = self.R
        cl, cr = self.cl, self.cr
        if rl > 0 and s < ul - cl:
            return self.L
        if rr > 0 and s >= ur + cr:
            return self.R
        if rl > 0 and s < ul + 2 * cl / (g - 1):
            c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
            return (rl * (c / cl) ** (2 / (g - 1)), 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s), pl * (c / cl) ** (2 * g / (g - 1)))
        if rr > 0 and s > ur - 2 * cr / (g - 1):
            c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
            return (rr * (c / cr) ** (2 / (g - 1)), 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s), pr * (c / cr) ** (2 * g / (g - 1)))
        return (0.0, 0.0, 0.0)

    def sample(self, s):
        if self.vacuum:
            return self._vacuum_sample(s)
        g, p, u = self.g, self.p, self.u
        if s <= u:
            rk, uk, pk = self.L
            c = self.cl
            if p > pk:
                S = uk - c * math.sqrt((g + 1) / (2 * g) * p / pk + (g - 1) / (2 * g))
                if s <= S:
                    return self.L
                r, k = p / pk, (g - 1) / (g + 1)
                return (rk * (r + k) / (k * r + 1), u, p)
            head = uk - c
            tail = u - c * (p / pk) ** ((g - 1) / (2 * g))
            if s <= head:
                return self.L
            if s >= tail:
                return (rk * (p / pk) ** (1 / g), u, p)
            cf = 2 / (g + 1) * (c + (g - 1) / 2 * (uk - s))
            return (rk * (cf / c) ** (2 / (g - 1)), 2 / (g + 1) * (c + (g - 1) / 2 * uk + s),
                    pk * (cf / c) ** (2 * g / (g - 1)))
        rk, uk, pk = self.R
        c = self.cr
        if p > pk:
            S = uk + c * math.sqrt((g + 1) / (2 * g) * p / pk + (g - 1) / (2 * g))
            if s >= S:
                return self.R
            r, k = p / pk, (g - 1) / (g + 1)
            return (rk * (r + k) / (k * r + 1), u, p)
        head = uk + c
        tail = u + c * (p / pk) ** ((g - 1) / (2 * g))
        if s >= head:
            return self.R
        if s <= tail:
            return (rk * (p / pk) ** (1 / g), u, p)
        cf = 2 / (g + 1) * (c - (g - 1) / 2 * (uk - s))
        return (rk * (cf / c) ** (2 / (g - 1)), 2 / (g + 1) * (-c + (g - 1) / 2 * uk + s),
                pk * (cf / c) ** (2 * g / (g - 1)))

    def average(self, a, b, x0, t, samples=256):
        """Conservative average over [a, b] at time t, jump at x0, as primitives (rho, u, p)."""
        g = self.g
        m = mom = e = 0.0
        for k in range(samples):
            x = a + (b - a) * (k + 0.5) / samples
            r, u, p = self.sample((x - x0) / t)
            m += r
            mom += r * u
            e += p / (g - 1) + 0.5 * r * u * u
        m /= samples
        mom /= samples
        e /= samples
        u = mom / m if m > 0 else 0.0
        return (m, u, (g - 1) * (e - 0.5 * m * u * u))
