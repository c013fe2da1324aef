"""ts_host.py - the oracle's restatement of the spin-temperature host tables.

TEST INFRASTRUCTURE ONLY: imported by tests/ and oracle.py, always as the
checker; the product never imports it.  It shares no code with the library:
every function below is restated here from the reference, in numpy/Python,
and the tests pin it against scipy and closed forms.

Reference functions restated (src/core/...):
  table files and splines     XRayHeatingFunctions.c:105-223 (RECFAST),
                              :1602-1845 (kappa_10, kappa_10_elec, kappa_10_pH),
                              :586-683 (stellar spectra), :805-912 (x_int tables)
  x_int interpolation         :914-1220 (interp_fheat, interp_n_Lya, interp_nion_*,
                              locate_energy_index, locate_xHII_index)
  cross sections              :409-467
  cosmology                   :377-404 (dtdz, hubble), :470-474 (drdz),
                              :1469-1511 (dicke, ddicke_dz), :1884-1917 (interpolate_fcoll)
  tauX, nu_tau_one            :228-375
  frequency integrals         :685-803 (integrand_in_nu_*, integrate_over_nu)
  Ly-n sum helpers            :496-584 (zmax, nu_n, frecycle), :586-683 (spectral_emissivity)
  shell plan                  ComputeTs.c:276-377 (radii, light-cone edges, SFR-history
                              operations; load_reion_sfr_grids reionization.c:1756-1784)
  snapshot tables             ComputeTs.c:517-716 (freq_int_*_tbl, sum_lyn and its kink
                              fix, growth factor, const_zp_prefactor)
  evolveInt's shell factors   XRayHeatingFunctions.c:1289-1311

GSL is absent here; its three routines are restated from their published
algorithms: the natural cubic spline (gsl_interp_cspline), QAG with the
61-point Gauss-Kronrod rule (QUADPACK's qag, Piessens et al. 1983: bisect the
interval of largest error until the summed error meets epsrel |result|), and
Brent's root bracketing (Brent 1973, in GSL's form with its interval test).
The 61-point rule is computed here from its definition (Gauss-Legendre 30
nodes plus the 31 roots of the Stieltjes polynomial), not taken from a table.
C float arithmetic (the x_int tables, float arguments of dtdz/hubble/zmax)
is reproduced with numpy.float32 operations.
"""
import math
import os

import numpy as np

F32 = np.float32

# constants: src/meraxes.h:21-44, reionization.h:11-42, XRayHeatingFunctions.h:13-39, ComputeTs.h:6
PLANCK = 6.6262e-27
HUBBLE = 3.2407789e-18
GRAVITY = 6.672e-8
PROTONMASS = 1.6726e-24
SOLAR_MASS = 1.989e33
SPEED_OF_LIGHT = 2.9979e10
SEC_PER_YEAR = 3.155e7
MPC = 3.086e24
TCMB = 2.728
SIGMA_HI = 6.3e-18
TINY = 1e-30
L_FACTOR = 0.620350491
R_XLY_MAX = float(F32(500))
LY_ALPHA_HZ = 2.46606727e15
NU_OVER_EV = 1.60217646e-12 / PLANCK
NUIONIZATION = 13.60 * NU_OVER_EV
HEI_NUIONIZATION = 24.59 * NU_OVER_EV
HEII_NUIONIZATION = NUIONIZATION * 4
NSPEC_MAX = 23
NXHII = 14
NENERGY = 258

# x_int_XHII, XRayHeatingFunctions.c:822-836 (a float array)
XHII = np.array([1.0e-4, 2.318e-4, 4.677e-4, 1.0e-3, 2.318e-3, 4.677e-3, 1.0e-2, 2.318e-2, 4.677e-2, 1.0e-1,
                 0.5, 0.9, 0.99, 0.999], dtype=F32)

# kappa_10 table, Zygelman (2005) Table 2 column 4 (XRayHeatingFunctions.c:1611-1664): data
_KAPPA_HH = ((1.0, 1.38e-13), (2.0, 1.43e-13), (4.0, 2.71e-13), (6.0, 6.60e-13), (8.0, 1.47e-12),
             (10.0, 2.88e-12), (15.0, 9.10e-12), (20.0, 1.78e-11), (25.0, 2.73e-11), (30.0, 3.67e-11),
             (40.0, 5.38e-11), (50.0, 6.86e-11), (60.0, 8.14e-11), (70.0, 9.25e-11), (80.0, 1.02e-10),
             (90.0, 1.11e-10), (100.0, 1.19e-10), (200.0, 1.75e-10), (300.0, 2.09e-10), (501.0, 2.565e-10),
             (701.0, 2.91e-10), (1000.0, 3.31e-10), (2000.0, 4.27e-10), (3000.0, 4.97e-10),
             (5000.0, 6.03e-10), (7000.0, 6.87e-10), (10000.0, 7.87e-10))

# recycling fractions of the Ly-n resonances (XRayHeatingFunctions.c:515-584): data
_FRECYCLE = (1, 1, 1, 0, 0.2609, 0.3078, 0.3259, 0.3353, 0.3410, 0.3448, 0.3476, 0.3496, 0.3512, 0.3524,
             0.3535, 0.3543, 0.3550, 0.3556, 0.3561, 0.3565, 0.3569, 0.3572, 0.3575, 0.3578, 0.3580, 0.3582,
             0.3584, 0.3586, 0.3587, 0.3589, 0.3590)


def frecycle(n):
    return float(_FRECYCLE[n]) if 0 <= n <= 30 else 0.0


# ---------------------------------------------------------------- splines

class NaturalSpline:
    """Natural cubic spline through (x, y), evaluated as gsl_interp_cspline does:
    the interval from a bisection search, then y_i + dx (b_i + dx (c_i + dx d_i))
    with b_i, d_i formed from the second-derivative coefficients c."""

    def __init__(self, x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        n = len(x)
        h = np.diff(x)
        c = np.zeros(n)
        if n > 2:
            # tridiagonal system for c_1..c_{n-2}, c_0 = c_{n-1} = 0
            m = n - 2
            diag = 2.0 * (h[:-1] + h[1:])
            off = h[1:-1]
            rhs = 3.0 * ((y[2:] - y[1:-1]) / h[1:] - (y[1:-1] - y[:-2]) / h[:-1])
            # Thomas algorithm (symmetric, diagonally dominant)
            cp = np.zeros(m)
            dp = np.zeros(m)
            for i in range(m):
                a = off[i - 1] if i > 0 else 0.0
                den = diag[i] - (a * cp[i - 1] if i > 0 else 0.0)
                cp[i] = (off[i] / den) if i < m - 1 else 0.0
                dp[i] = (rhs[i] - (a * dp[i - 1] if i > 0 else 0.0)) / den
            sol = np.zeros(m)
            for i in range(m - 1, -1, -1):
                sol[i] = dp[i] - (cp[i] * sol[i + 1] if i < m - 1 else 0.0)
            c[1:-1] = sol
        self.x, self.y, self.c = x, y, c

    def index(self, xv):
        lo, hi = 0, len(self.x) - 1
        while hi > lo + 1:
            i = (hi + lo) // 2
            if self.x[i] > xv:
                hi = i
            else:
                lo = i
        return lo

    def __call__(self, xv):
        xv = float(xv)
        i = self.index(xv)
        x, y, c = self.x, self.y, self.c
        h = x[i + 1] - x[i]
        dx = xv - x[i]
        dy = y[i + 1] - y[i]
        b = dy / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0
        d = (c[i + 1] - c[i]) / (3.0 * h)
        return y[i] + dx * (b + dx * (c[i] + dx * d))


# ------------------------------------------------- quadrature and roots

def _legendre_all(x, m):
    """P_0..P_m at the points x (three-term recurrence)."""
    x = np.asarray(x, np.float64)
    p = [np.ones_like(x), x.copy()]
    for k in range(1, m):
        p.append(((2 * k + 1) * x * p[k] - k * p[k - 1]) / (k + 1))
    return p


def _gauss_kronrod_61():
    """Nonnegative nodes (descending, Kronrod nodes at even positions, 0 last),
    their Kronrod weights and the Gauss weights of the odd positions."""
    from scipy.optimize import brentq

    gx, gw = np.polynomial.legendre.leggauss(30)
    pos = np.sort(gx[gx > 0])[::-1]
    gwpos = gw[gx > 0][np.argsort(gx[gx > 0])[::-1]]
    # Stieltjes polynomial E_31 = sum over odd j of c_j P_j, c_31 = 1, orthogonal
    # under the weight P_30 to every odd P_k, k < 31 (even k vanish by parity)
    qx, qw = np.polynomial.legendre.leggauss(64)  # exact to degree 127 >= 30 + 31 + 29
    Pq = _legendre_all(qx, 31)
    odd = list(range(1, 30, 2))
    A = np.array([[np.sum(qw * Pq[30] * Pq[j] * Pq[k]) for j in odd] for k in odd])
    b = np.array([-np.sum(qw * Pq[30] * Pq[31] * Pq[k]) for k in odd])
    cj = np.linalg.solve(A, b)
    coef = np.zeros(32)
    coef[odd] = cj
    coef[31] = 1.0
    E = lambda t: np.polynomial.legendre.legval(t, coef)
    edges = [1.0] + list(pos)
    kr = [brentq(E, edges[i + 1], edges[i], xtol=1e-16, rtol=4 * np.finfo(float).eps) for i in range(15)]
    nodes = []
    for i in range(15):
        nodes += [kr[i], pos[i]]
    nodes.append(0.0)
    nodes = np.array(nodes)
    # weights: exact for the even Legendre polynomials P_0 .. P_60
    Pn = _legendre_all(nodes, 60)
    mult = np.where(nodes > 0, 2.0, 1.0)
    M = np.array([mult * Pn[2 * k] for k in range(31)])
    rhs = np.zeros(31)
    rhs[0] = 2.0
    wk = np.linalg.solve(M, rhs)
    return nodes, wk, gwpos


XGK, WGK, WG = _gauss_kronrod_61()
_EPS = np.finfo(float).eps
_DBL_MIN = np.finfo(float).tiny


def _rescale_error(err, result_abs, result_asc):
    err = abs(err)
    if result_asc != 0 and err != 0:
        scale = (200 * err / result_asc) ** 1.5
        err = result_asc * scale if scale < 1 else result_asc
    if result_abs > _DBL_MIN / (50 * _EPS):
        min_err = 50 * _EPS * result_abs
        if min_err > err:
            err = min_err
    return err


def gk61(f, a, b):
    """61-point Gauss-Kronrod on [a, b]: (result, abserr, resabs, resasc).
    f maps an array of abscissae to an array of values; the sums are taken in
    the rule's own order (centre, then the Gauss pairs, then the others)."""
    center = 0.5 * (a + b)
    half = 0.5 * (b - a)
    ab = half * XGK[:30]
    xs = np.concatenate(([center], center - ab, center + ab))
    fv = np.asarray(f(xs), np.float64)
    fc = float(fv[0])
    f1 = fv[1:31]
    f2 = fv[31:61]
    res_g = 0.0
    res_k = fc * float(WGK[30])
    res_abs = abs(res_k)
    for j in range(15):
        jt = 2 * j + 1
        s = float(f1[jt]) + float(f2[jt])
        res_g += float(WG[j]) * s
        res_k += float(WGK[jt]) * s
        res_abs += float(WGK[jt]) * (abs(float(f1[jt])) + abs(float(f2[jt])))
    for j in range(15):
        jt = 2 * j
        res_k += float(WGK[jt]) * (float(f1[jt]) + float(f2[jt]))
        res_abs += float(WGK[jt]) * (abs(float(f1[jt])) + abs(float(f2[jt])))
    mean = res_k * 0.5
    res_asc = float(WGK[30]) * abs(fc - mean)
    for j in range(30):
        res_asc += float(WGK[j]) * (abs(float(f1[j]) - mean) + abs(float(f2[j]) - mean))
    err = (res_k - res_g) * half
    res_k *= half
    res_abs *= abs(half)
    res_asc *= abs(half)
    return res_k, _rescale_error(err, res_abs, res_asc), res_abs, res_asc


def qag61(f, a, b, epsrel, limit=1000):
    """QUADPACK's QAG with absolute tolerance 0 (every call site in the
    reference passes epsabs = 0).  Returns (result, abserr, n_intervals)."""
    r0, e0, resabs, resasc = gk61(f, a, b)
    tol = epsrel * abs(r0)
    round_off = 50 * _EPS * resabs
    if (e0 <= round_off and e0 > tol) or (e0 <= tol and e0 != resasc) or e0 == 0.0 or limit == 1:
        return r0, e0, 1
    # interval list in storage order (QUADPACK's: the bisected interval keeps its
    # slot for the half with the smaller error, the other half is appended)
    al, bl, rl, el = [a], [b], [r0], [e0]
    area, errsum = r0, e0
    rt1 = rt2 = 0
    it = 1
    while True:
        k = max(range(len(el)), key=lambda i: (el[i], -i))
        a1, b1 = al[k], 0.5 * (al[k] + bl[k])
        a2, b2 = b1, bl[k]
        r1, e1, _, s1 = gk61(f, a1, b1)
        r2, e2, _, s2 = gk61(f, a2, b2)
        area12, error12 = r1 + r2, e1 + e2
        errsum += error12 - el[k]
        area += area12 - rl[k]
        if s1 != e1 and s2 != e2:
            if abs(rl[k] - area12) <= 1.0e-5 * abs(area12) and error12 >= 0.99 * el[k]:
                rt1 += 1
            if it >= 10 and error12 > el[k]:
                rt2 += 1
        tol = epsrel * abs(area)
        stop = False
        if errsum > tol:
            if rt1 >= 6 or rt2 >= 20:
                stop = True
            if max(abs(a1), abs(b2)) <= (1 + 100 * _EPS) * (abs(a2) + 1000 * _DBL_MIN):
                stop = True
        if e2 > e1:
            al[k], rl[k], el[k] = a2, r2, e2
            al.append(a1)
            bl.append(b1)
            rl.append(r1)
            el.append(e1)
        else:
            bl[k], rl[k], el[k] = b1, r1, e1
            al.append(a2)
            bl.append(b2)
            rl.append(r2)
            el.append(e2)
        it += 1
        if stop or it >= limit or errsum <= tol:
            break
    result = 0.0
    for r in rl:
        result += r
    return result, errsum, len(rl)


def brent_gsl(f, lo, hi, relerr, max_iter=100):
    """Brent's method as GSL's brent solver runs it, iterated until
    |x_upper - x_lower| < relerr * min(|x_lower|, |x_upper|) (the bounds of
    each iteration: the new point and the end of opposite sign).  Returns
    (root, iterations); raises when the end points do not straddle 0."""
    a, b = lo, hi
    fa, fb = f(lo), f(hi)
    if (fa < 0.0 and fb < 0.0) or (fa > 0.0 and fb > 0.0):
        raise ValueError("endpoints do not straddle y=0")
    c, fc = b, fb
    d = e = b - a
    r = 0.5 * (lo + hi)
    for it in range(1, max_iter + 1):
        ac_equal = False
        if (fb < 0 and fc < 0) or (fb > 0 and fc > 0):
            ac_equal = True
            c, fc = a, fa
            d = e = b - a
        if abs(fc) < abs(fb):
            ac_equal = True
            a, b, c = b, c, b
            fa, fb, fc = fb, fc, fb
        tol = 0.5 * _EPS * abs(b)
        m = 0.5 * (c - b)
        if fb == 0:
            return b, it
        if abs(m) <= tol:
            return b, it
        if abs(e) < tol or abs(fa) <= abs(fb):
            d = e = m
        else:
            s = fb / fa
            if ac_equal:
                p = 2 * m * s
                q = 1 - s
            else:
                q = fa / fc
                rr = fb / fc
                p = s * (2 * m * q * (q - rr) - (b - a) * (rr - 1))
                q = (q - 1) * (rr - 1) * (s - 1)
            if p > 0:
                q = -q
            else:
                p = -p
            if 2 * p < min(3 * m * q - abs(tol * q), abs(e * q)):
                e = d
                d = p / q
            else:
                d = e = m
        a, fa = b, fb
        if abs(d) > tol:
            b += d
        else:
            b += tol if m > 0 else -tol
        fb = f(b)
        r = b
        cc = a if ((fb < 0 and fc < 0) or (fb > 0 and fc > 0)) else c
        xl, xu = (b, cc) if b < cc else (cc, b)
        min_abs = min(abs(xl), abs(xu)) if ((xl > 0 and xu > 0) or (xl < 0 and xu < 0)) else 0.0
        if abs(xu - xl) < relerr * min_abs:
            return r, it
    return r, max_iter


# ------------------------------------------------------------- the data

def _f32(tok):
    return float(F32(float(tok)))


class TsData:
    """The 21cmFAST tables of TablesForXHeatingDir, read as the reference reads
    them (float fields, converted to double where the reference does)."""

    def __init__(self, tables_dir, nion_phot_per_bary):
        self.dir = tables_dir
        # RECFAST: rows from z = 500 down, stored in ascending z (:128-133, :189-193)
        rows = [l.split() for l in open(os.path.join(tables_dir, "recfast_LCDM.dat")) if l.strip()]
        rows = rows[:501][::-1]
        zt = [_f32(r[0]) for r in rows]
        self.recfast_zmax = zt[-1]
        self.recfast_T = NaturalSpline(zt, [_f32(r[3]) for r in rows])
        self.recfast_xe = NaturalSpline(zt, [_f32(r[1]) for r in rows])
        # kappa splines in log T -> log kappa
        tk = [math.log(t) for t, _ in _KAPPA_HH]
        kp = [math.log(k) for _, k in _KAPPA_HH]
        self.kappa_HH = NaturalSpline(tk, kp)
        self.kappa_eH = self._kappa_file("kappa_eH_table.dat", 20)
        self.kappa_pH = self._kappa_file("kappa_pH_table.dat", 17)
        self._read_spectra(nion_phot_per_bary)
        self._read_xint()

    def _kappa_file(self, name, n):
        rows = [l.split() for l in open(os.path.join(self.dir, name)) if l.strip()][:n]
        return NaturalSpline([math.log(_f32(r[0])) for r in rows], [math.log(_f32(r[1])) for r in rows])

    def _read_spectra(self, nion):
        # spectral_emissivity flag 1 (:617-641): float table, Pop II normalisation
        rows = [l.split() for l in open(os.path.join(self.dir, "stellar_spectra.dat")) if l.strip()]
        n = [0] * NSPEC_MAX
        N0 = [F32(0)] * NSPEC_MAX
        al = [F32(0)] * NSPEC_MAX
        for i in range(1, NSPEC_MAX):
            r = rows[i - 1]
            n[i] = int(r[0])
            N0[i] = F32(float(r[1]))
            al[i] = F32(float(r[2]))
        nu = [F32(0)] * NSPEC_MAX
        for i in range(1, NSPEC_MAX):
            nu[i] = F32(4.0 / 3.0 * (1.0 - 1.0 / math.pow(n[i], 2.0)))
        for i in range(1, NSPEC_MAX - 1):
            a1 = float(al[i] + F32(1))  # float + int: float arithmetic
            fac = math.pow(float(nu[i + 1]), a1) - math.pow(float(nu[i]), a1)
            N0[i] = F32(float(N0[i]) * (a1 / fac * nion))
        self.spec_nu = [float(v) for v in nu]
        self.spec_N0 = [float(v) for v in N0]
        self.spec_alpha = [float(v) for v in al]

    def _read_xint(self):
        # initialize_interp_arrays (:805-912): 14 files, 258 energies each
        tabs = {k: np.zeros((NXHII, NENERGY), F32) for k in ("fheat", "nLya", "nHI", "nHeI", "nHeII")}
        E = None
        for m in range(NXHII):
            x = float(XHII[m])
            if x < 0.3:
                name = "log_xi_%1.1f.dat" % math.log10(x)
            else:
                name = "xi_%1.3f.dat" % x
            tok = open(os.path.join(self.dir, "x_int_tables", name)).read().split()
            vals = tok[5 + 5 + 11:]  # 5 labels, 5 numbers, 11 column headings
            arr = np.array([float(v) for v in vals[:9 * NENERGY]], np.float64).astype(F32).reshape(NENERGY, 9)
            E = arr[:, 0]
            tabs["fheat"][m] = arr[:, 2]
            tabs["nLya"][m] = arr[:, 4]
            tabs["nHI"][m] = arr[:, 5]
            tabs["nHeI"][m] = arr[:, 6]
            tabs["nHeII"][m] = arr[:, 7]
        self.E = E
        self.xint = tabs

    # -- spectral_emissivity, flag 0, Pop II (:642-660)
    def spectral_emissivity(self, nu_norm):
        for i in range(1, NSPEC_MAX - 1):
            if nu_norm >= self.spec_nu[i] and nu_norm < self.spec_nu[i + 1]:
                return self.spec_N0[i] * math.pow(nu_norm, self.spec_alpha[i]) / LY_ALPHA_HZ
        i = NSPEC_MAX - 1  # below Ly-alpha or above the last edge: the last segment's power law
        return self.spec_N0[i] * math.pow(nu_norm, self.spec_alpha[i]) / LY_ALPHA_HZ

    # -- interp_fheat / interp_n_Lya / interp_nion_* (:914-1192), vectorised over En
    def interp(self, which, En, xHII):
        tab = self.xint[which]
        E = self.E
        En = np.asarray(En, F32).copy()
        top = float(E[NENERGY - 1])
        En = np.where(En.astype(np.float64) > 0.999 * top, F32(top * 0.999), En)
        below = En < E[0]
        x = F32(xHII)
        if float(x) > float(XHII[NXHII - 1]) * 0.999:
            x = F32(float(XHII[NXHII - 1]) * 0.999)
        elif x < XHII[0]:
            x = F32(1.001 * float(XHII[0]))
        Es = np.where(below, E[0], En)
        Ed = Es.astype(np.float64)
        n_lo = np.where(Ed < 1008.88, (np.log(Ed / 10.0) / 1.98026273e-2).astype(np.int64),
                        232 + (np.log(Ed / 1008.88) / 9.53101798e-2).astype(np.int64))
        n_hi = n_lo + 1
        m_lo = NXHII - 1
        while x < XHII[m_lo]:
            m_lo -= 1
        m_hi = m_lo + 1
        dE = E[n_hi] - E[n_lo]
        off = Es - E[n_lo]
        elow = (tab[m_lo, n_hi] - tab[m_lo, n_lo]) / dE
        elow = elow * off
        elow = elow + tab[m_lo, n_lo]
        ehigh = (tab[m_hi, n_hi] - tab[m_hi, n_lo]) / dE
        ehigh = ehigh * off
        ehigh = ehigh + tab[m_hi, n_lo]
        res = (ehigh - elow) / (XHII[m_hi] - XHII[m_lo])
        res = res * (x - XHII[m_lo])
        res = res + elow
        return np.where(below, F32(1.0 if which == "fheat" else 0.0), res).astype(F32)


# ------------------------------------------------------ scalar physics

def HI_ion_crosssec(nu):
    nu = np.asarray(nu, np.float64)
    nu = np.where(nu == NUIONIZATION, nu + TINY, nu)
    ok = nu >= NUIONIZATION
    nus = np.where(ok, nu, 2 * NUIONIZATION)
    eps = np.sqrt(nus / NUIONIZATION - 1)
    v = (SIGMA_HI / 1.0 / 1.0 * (NUIONIZATION / nus) ** 4 * np.exp(4 - (4 * np.arctan(eps) / eps)) /
         (1 - np.exp(-2 * math.pi / eps)))
    return np.where(ok, v, 0.0)


def HeII_ion_crosssec(nu):
    nu = np.asarray(nu, np.float64)
    nu = np.where(nu == HEII_NUIONIZATION, nu + TINY, nu)
    ok = nu >= HEII_NUIONIZATION
    nus = np.where(ok, nu, 2 * HEII_NUIONIZATION)
    eps = np.sqrt(nus / HEII_NUIONIZATION - 1)
    v = ((6.3e-18) / 2.0 / 2.0 * (HEII_NUIONIZATION / nus) ** 4 * np.exp(4 - (4 * np.arctan(eps) / eps)) /
         (1 - np.exp(-2 * math.pi / eps)))
    return np.where(ok, v, 0.0)


def HeI_ion_crosssec(nu):
    nu = np.asarray(nu, np.float64)
    ok = nu >= HEI_NUIONIZATION
    x = nu / NU_OVER_EV / 13.61 - 0.4434
    y = np.sqrt(x * x + 2.136 ** 2)
    v = 9.492e-16 * ((x - 1) * (x - 1) + 2.039 * 2.039) * y ** (0.5 * 3.188 - 5.5) * (1.0 + np.sqrt(y / 1.469)) ** -3.188
    return np.where(ok, v, 0.0)


class TsSnapshot:
    """Plan and tables of one snapshot: what _ComputeTs computes on the host."""


class TsHost:
    """Cosmology and physics of one parameter set (an mrx_ts_params)."""

    def __init__(self, p, data=None):
        self.p = p
        tables_dir = p.tables_dir.decode() if isinstance(p.tables_dir, bytes) else p.tables_dir
        self.d = data if data is not None else TsData(tables_dir, p.nion_phot_per_bary)
        h = p.hubble_h
        rhocrit_cgs = 3.0 * HUBBLE * HUBBLE * h * h / (8.0 * math.pi * GRAVITY)
        OMb = p.baryon_frac * p.omega_m
        self.No = rhocrit_cgs * OMb * (1 - p.y_he) / PROTONMASS
        self.He_No = rhocrit_cgs * OMb * p.y_he / (4.0 * PROTONMASS)
        self.f_H = self.No / (self.No + self.He_No)
        self.f_He = self.He_No / (self.No + self.He_No)
        self.N_b0 = self.No + self.He_No

    # -- dtdz(float z) (:377-392): 1 + z is float arithmetic
    def dtdz(self, z):
        p = self.p
        opz = float(F32(1) + F32(z))
        OMl, OMm = p.omega_lambda, p.omega_m
        x = math.sqrt(OMl / OMm) * math.pow(opz, -3.0 / 2.0)
        dxdz = math.sqrt(OMl / OMm) * math.pow(opz, -5.0 / 2.0) * (-3.0 / 2.0)
        const1 = 2 * math.sqrt(1 + OMm / OMl) / (3.0 * HUBBLE * p.hubble_h)
        numer = dxdz * (1 + x * math.pow(math.pow(x, 2) + 1, -0.5))
        denom = x + math.sqrt(math.pow(x, 2) + 1)
        return const1 * numer / denom

    def dtdz_vec(self, z):
        p = self.p
        opz = (F32(1) + np.asarray(z, np.float64).astype(F32)).astype(np.float64)
        OMl, OMm = p.omega_lambda, p.omega_m
        x = math.sqrt(OMl / OMm) * opz ** (-3.0 / 2.0)
        dxdz = math.sqrt(OMl / OMm) * opz ** (-5.0 / 2.0) * (-3.0 / 2.0)
        const1 = 2 * math.sqrt(1 + OMm / OMl) / (3.0 * HUBBLE * p.hubble_h)
        numer = dxdz * (1 + x * (x ** 2 + 1) ** -0.5)
        denom = x + np.sqrt(x ** 2 + 1)
        return const1 * numer / denom

    # -- hubble(float z) (:394-404)
    def hubble(self, z):
        p = self.p
        opz = float(F32(1) + F32(z))
        return HUBBLE * p.hubble_h * math.sqrt(p.omega_m * math.pow(opz, 3) + p.omega_r * math.pow(opz, 4) +
                                               p.omega_lambda)

    # -- drdz(float z) (:470-474): here 1.0 + z is double arithmetic
    def drdz(self, z):
        return (1.0 + float(F32(z))) * SPEED_OF_LIGHT * self.dtdz(z)

    # -- dicke (:1469-1504), ddicke_dz (:1506-1511)
    def dicke(self, z):
        p = self.p
        tiny = 1e-4
        OMm, OMl, OMr, OMk, wl = p.omega_m, p.omega_lambda, p.omega_r, p.omega_k, p.w_lambda
        OMtot = OMm + OMl + OMr + OMk
        if abs(OMm - 1.0) < tiny:
            return 1.0 / (1.0 + z)
        if OMl > -tiny and abs(OMl + OMm + OMr - 1.0) < 0.01 and abs(wl + 1.0) < tiny:
            om_z = OMm * math.pow(1 + z, 3) / (OMl + OMm * math.pow(1 + z, 3) + OMr * math.pow(1 + z, 4))
            dz = 2.5 * om_z / (1.0 / 70.0 + om_z * (209 - om_z) / 140.0 + math.pow(om_z, 4.0 / 7.0))
            d0 = 2.5 * OMm / (1.0 / 70.0 + OMm * (209 - OMm) / 140.0 + math.pow(OMm, 4.0 / 7.0))
            return dz / (d0 * (1.0 + z))
        if OMtot < 1 + tiny and abs(OMl) < tiny:
            x0 = 1.0 / (OMm + 0.0) - 1.0
            d0 = 1 + 3.0 / x0 + 3 * math.log(math.sqrt(1 + x0) - math.sqrt(x0)) * math.sqrt(1 + x0) / x0 ** 1.5
            x = abs(1.0 / (OMm + 0.0) - 1.0) / (1 + z)
            dz = 1 + 3.0 / x + 3 * math.log(math.sqrt(1 + x) - math.sqrt(x)) * math.sqrt(1 + x) / x ** 1.5
            return dz / d0
        return -1.0

    def ddicke_dz(self, z):
        dz = float(F32(1e-10))
        return (self.dicke(z + dz) - self.dicke(z)) / dz

    def T_RECFAST(self, z):
        return self.d.recfast_T(float(F32(z)))

    def xion_RECFAST(self, z):
        return self.d.recfast_xe(float(F32(z)))

    # -- interpolate_fcoll, Pop II (:1884-1900), vectorised
    def interpolate_fcoll(self, redshift, stored_fcoll):
        s = self.p.snapshot
        ZZ = self.ZZ
        F = stored_fcoll
        if s == 0:
            v = np.full_like(np.asarray(redshift, np.float64), F[0])
        else:
            v = F[s - 1] + (np.asarray(redshift, np.float64) - ZZ[s - 1]) * (F[s] - F[s - 1]) / (ZZ[s] - ZZ[s - 1])
        return np.where(v < 0.0, 0.0, v)

    def species_xsec(self, nu, x_e):
        return (self.f_H * (1 - x_e) * HI_ion_crosssec(nu) + self.f_He * (1 - x_e) * HeI_ion_crosssec(nu) +
                self.f_He * x_e * HeII_ion_crosssec(nu))

    # -- tauX (:344-375) with tauX_integrand (:310-342), Pop II only
    def tauX(self, nu, x_e, x_e_ave, zp, zpp, stored_fcoll):
        nu_0 = nu / (1 + zp)
        eff = self.p.reion_efficiency

        def integrand(zhat):
            drpropdz = SPEED_OF_LIGHT * self.dtdz_vec(zhat)
            n = self.N_b0 * (1 + zhat) ** 3
            nuhat = nu_0 * (1 + zhat)
            fcoll = self.interpolate_fcoll(zhat, stored_fcoll)
            hff = np.where(fcoll < 1e-20, 1.0, 1 - (eff * fcoll) / (1.0 - x_e_ave))
            hff = np.where(hff < 1e-4, 1e-4, hff)
            return drpropdz * n * hff * self.species_xsec(nuhat, x_e)

        return qag61(integrand, zpp, zp, 0.005)[0]

    # -- nu_tau_one (:228-307)
    def nu_tau_one(self, zp, zpp, x_e, x_e_ave, stored_fcoll):
        if x_e > 0.9999:
            return -1.0
        if self.tauX(HEI_NUIONIZATION, x_e, x_e_ave, zp, zpp, stored_fcoll) < 1:
            return HEI_NUIONIZATION
        f = lambda nu: self.tauX(nu, x_e, x_e_ave, zp, zpp, stored_fcoll) - 1
        return brent_gsl(f, HEI_NUIONIZATION, 1e6 * NU_OVER_EV, 0.02, 100)[0]

    # -- integrand_in_nu_*_integral (:685-753), vectorised over nu; each species'
    # term multiplied left to right as there, the n_ion sums in float
    def nu_integrand(self, flag, x_e):
        p = self.p
        thr = p.nu_xray_gal_threshold
        alpha = p.spec_index_xray_gal
        d = self.d
        xf = F32(x_e)
        species = ((NUIONIZATION, self.f_H, 1 - x_e, HI_ion_crosssec),
                   (HEI_NUIONIZATION, self.f_He, 1 - x_e, HeI_ion_crosssec),
                   (HEII_NUIONIZATION, self.f_He, x_e, HeII_ion_crosssec))

        def f(nu):
            nu = np.asarray(nu, np.float64)
            total = None
            for nu_ion, frac, xfac, xsec in species:
                En = ((nu - nu_ion) / NU_OVER_EV).astype(F32)
                if flag == 0:
                    term = d.interp("fheat", En, xf).astype(np.float64) * PLANCK * (nu - nu_ion)
                elif flag == 1:
                    Fi = d.interp("nHI", En, xf) + d.interp("nHeI", En, xf) + d.interp("nHeII", En, xf) + F32(1)
                    term = Fi.astype(np.float64)
                else:
                    term = d.interp("nLya", En, xf).astype(np.float64)
                term = term * frac * xfac * xsec(nu)
                total = term if total is None else total + term
            return total * (nu / (thr * NU_OVER_EV)) ** (-alpha - 1)

        return f

    def integrate_over_nu(self, zp, local_x_e, lower, flag):
        r = qag61(self.nu_integrand(flag, local_x_e), lower, self.p.nu_xray_max * NU_OVER_EV, 0.01)[0]
        if flag == 2:
            return r * SPEED_OF_LIGHT / (4 * math.pi) / LY_ALPHA_HZ / self.hubble(zp)
        return r

    # -- the shell plan (ComputeTs.c:276-377)
    def plan(self):
        p = self.p
        snapshot = p.snapshot
        ZZ = [float(p.redshifts[i]) for i in range(snapshot + 1)]
        self.ZZ = ZZ
        s = TsSnapshot()
        s.nsteps = ns = p.num_filter_steps
        zp = ZZ[snapshot]
        prev_redshift = ZZ[snapshot] if snapshot == 0 else ZZ[snapshot - 1]
        s.zp = zp
        s.dzp = zp - prev_redshift
        s.dt_dzp = self.dtdz(float(F32(F32(zp))))
        s.homogeneous = (zp - p.max_heating_redshift) >= -0.0001
        s.ops = []  # (shell, kind, slot, weight): kinds 0 zero half, 1 scale half, 2 load, 3 add
        if s.homogeneous:
            return s
        box_size = p.box_size / p.hubble_h
        N = p.grid_dim
        R = L_FACTOR * box_size / float(F32(N))
        R_factor = math.pow(R_XLY_MAX / R, float(F32(1) / F32(ns)))  # 1 / (float)n: float division
        s.R = [0.0] * ns
        s.zpp_edge = [0.0] * ns
        scb = [0] * ns
        scb[0] = 1
        weight = 1.0
        for ct in range(ns):
            s.R[ct] = R
            if ct == 0:
                prev_zpp, prev_R = zp, 0.0
            else:
                prev_zpp, prev_R = s.zpp_edge[ct - 1], s.R[ct - 1]
                scb[ct] = scb[ct - 1]
            s.zpp_edge[ct] = prev_zpp - (R - prev_R) * MPC / self.drdz(prev_zpp)
            zpp = (s.zpp_edge[ct] + prev_zpp) * 0.5
            total_weight = (prev_zpp - s.zpp_edge[ct]) * self.dtdz(zpp)
            if snapshot - scb[ct] >= 0:
                if snapshot - scb[ct] == 0 and weight < 0:
                    s.ops.append((ct, 0, 0, 0.0))
                    scb[ct] += 1
                elif s.zpp_edge[ct] > ZZ[snapshot - scb[ct]]:
                    if ct == 0:
                        weight = s.dt_dzp * s.dzp / total_weight
                        s.ops.append((ct, 1, 0, weight))
                    elif ZZ[snapshot - scb[ct]] > s.zpp_edge[ct - 1]:
                        zs = ZZ[snapshot - scb[ct]]
                        weight = -self.dtdz(0.5 * (zs + s.zpp_edge[ct - 1]))
                        weight *= zs - s.zpp_edge[ct - 1]
                        weight /= total_weight
                        s.ops.append((ct, 2, scb[ct] - 1, weight))
                    else:
                        s.ops.append((ct, 0, 0, 0.0))
                    while s.zpp_edge[ct] > ZZ[snapshot - scb[ct]]:
                        if snapshot - scb[ct] == 0:
                            weight = -1.0
                            break
                        zs = ZZ[snapshot - scb[ct]]
                        if s.zpp_edge[ct] < ZZ[snapshot - scb[ct] - 1]:
                            # :345 - the ternary tests zpp_edge[R_ct-1] itself (non-zero), so
                            # zedge is the snapshot's redshift; zpp_edge[-1] at R_ct = 0 is
                            # read past the array there, taken as non-zero here
                            prev_edge = s.zpp_edge[ct - 1] if ct > 0 else 1.0
                            zedge = zs if prev_edge else float(prev_edge > zs)
                            weight = s.zpp_edge[ct] - zedge
                            weight *= -self.dtdz(0.5 * (s.zpp_edge[ct] + zedge))
                        else:
                            zo = ZZ[snapshot - scb[ct] - 1]
                            weight = zo - zs
                            weight *= -self.dtdz(0.5 * (zo + zs))
                        weight /= total_weight
                        s.ops.append((ct, 3, scb[ct], weight))
                        scb[ct] += 1
                elif ct > 0 and weight != 1:
                    weight = 1.0
                    s.ops.append((ct, 2, scb[ct] - 1, weight))
            R *= R_factor
        s.final_weight = weight
        # evolveInt's per-shell factors (XRayHeatingFunctions.c:1289-1311), zp a float there
        zpf = float(F32(zp))
        s.zpp = [0.0] * ns
        s.dzpp = [0.0] * ns
        for ct in range(ns):
            if ct == 0:
                s.zpp[ct] = (s.zpp_edge[0] + zpf) * 0.5
                s.dzpp[ct] = zpf - s.zpp_edge[0]
            else:
                s.zpp[ct] = (s.zpp_edge[ct] + s.zpp_edge[ct - 1]) * 0.5
                s.dzpp[ct] = s.zpp_edge[ct - 1] - s.zpp_edge[ct]
        s.dt_dzpp = [self.dtdz(z) for z in s.zpp]
        s.zpp_pow = [math.pow(1 + z, -p.spec_index_xray_gal) for z in s.zpp]
        return s

    # -- zmax (:496-503): float z, float arithmetic in 1 + z
    @staticmethod
    def zmax(z, n):
        num = 1 - math.pow(n + 1, -2)
        denom = 1 - math.pow(n, -2)
        return float(F32(float(F32(1) + F32(z)) * num / denom - 1))

    @staticmethod
    def nu_n(n):
        return (1.0 - math.pow(n, -2.0)) / 0.75

    def sum_lyn_raw(self, zp, zpp):
        total = 0.0
        for n in range(NSPEC_MAX, 1, -1):
            if zpp > self.zmax(zp, n):
                continue
            nuprime = self.nu_n(n) * (1 + zpp) / (1.0 + zp)
            total += frecycle(n) * self.d.spectral_emissivity(nuprime)
        return total

    # -- the snapshot tables (ComputeTs.c:517-716)
    def tables(self, s, collapse_fraction, x_e_ave, stored_fcoll=None):
        p = self.p
        if s.homogeneous:
            s.xe_recfast = self.xion_RECFAST(s.zp)
            s.T_recfast = self.T_RECFAST(s.zp)
            return s
        ns, zp = s.nsteps, s.zp
        fc = np.array([float(stored_fcoll[i]) for i in range(p.snapshot + 1)]) if stored_fcoll is not None \
            else np.zeros(p.snapshot + 1)
        fc[p.snapshot] = collapse_fraction
        s.stored_fcoll = fc
        s.collapse_fraction = collapse_fraction
        s.x_e_ave = x_e_ave
        s.no_light = 0 if collapse_fraction > 0.0 else 1
        s.nu_lower = [0.0] * ns
        s.sum_lyn = [0.0] * ns
        s.freq_heat = np.zeros((NXHII, ns))
        s.freq_ion = np.zeros((NXHII, ns))
        s.freq_lya = np.zeros((NXHII, ns))
        thr = p.nu_xray_gal_threshold * NU_OVER_EV
        first_radii = first_zero = True
        weight = s.final_weight
        for ct in range(ns):
            prev_zpp, prev_R = (zp, 0.0) if ct == 0 else (s.zpp_edge[ct - 1], s.R[ct - 1])
            zpp = (s.zpp_edge[ct] + prev_zpp) * 0.5
            lower = max(self.nu_tau_one(zp, zpp, x_e_ave, x_e_ave, fc), thr)
            s.nu_lower[ct] = lower
            for x in range(NXHII):
                xe = float(XHII[x])
                s.freq_heat[x, ct] = self.integrate_over_nu(zp, xe, lower, 0)
                s.freq_ion[x, ct] = self.integrate_over_nu(zp, xe, lower, 1)
                s.freq_lya[x, ct] = self.integrate_over_nu(zp, xe, lower, 2)
            s.sum_lyn[ct] = self.sum_lyn_raw(zp, zpp)
            # kink fix (:625-663)
            if ct > 2 and s.sum_lyn[ct] == 0.0 and s.sum_lyn[ct - 1] > 0.0 and first_radii:
                trial_max = (prev_zpp - (s.R[ct] - prev_R) * MPC / self.drdz(prev_zpp) + prev_zpp) * 0.5
                trial_min = (s.zpp_edge[ct - 2] - (s.R[ct - 1] - s.R[ct - 2]) * MPC / self.drdz(s.zpp_edge[ct - 2]) +
                             s.zpp_edge[ct - 2]) * 0.5
                for ii in range(1000):
                    trial = trial_min + (trial_max - trial_min) * float(F32(ii)) / (float(F32(1000)) - 1.)
                    counter = sum(1 for n in range(NSPEC_MAX, 1, -1) if not trial > self.zmax(zp, n))
                    if counter == 0 and first_zero:
                        first_zero = False
                        weight = float(F32(ii) / F32(1000))
                s.kink_shell = ct
                s.kink_weight = weight
                s.sum_lyn[ct] = weight * s.sum_lyn[ct - 1]
                first_radii = False
        s.growth_factor_zp = self.dicke(zp)
        s.dgrowth_factor_dzp = self.ddicke_dz(zp)
        s.dt_dzp = self.dtdz(zp)
        th, al = p.nu_xray_gal_threshold, p.spec_index_xray_gal
        if abs(al - 1.0) < 0.000001:
            conv = 1. / ((th * NU_OVER_EV) * math.log(p.nu_xray_soft_cut / th))
        else:
            conv = 1. / (math.pow(p.nu_xray_soft_cut * NU_OVER_EV, 1. - al) - math.pow(th * NU_OVER_EV, 1. - al))
            conv *= math.pow(th * NU_OVER_EV, -al) * (1 - al)
        conv *= SEC_PER_YEAR / PLANCK
        s.const_zp_prefactor = (p.l_xray_gal * conv) / (th * NU_OVER_EV) * SPEED_OF_LIGHT * math.pow(1 + zp, al + 3)
        return s


_DATA_CACHE = {}


def host(p):
    """A TsHost for the parameter set p, its table files read once per directory and N_ion."""
    key = (p.tables_dir, p.nion_phot_per_bary)
    if key not in _DATA_CACHE:
        d = p.tables_dir.decode() if isinstance(p.tables_dir, bytes) else p.tables_dir
        _DATA_CACHE[key] = TsData(d, p.nion_phot_per_bary)
    return TsHost(p, _DATA_CACHE[key])
