"""Pins the spin-temperature host tables: the oracle's numpy restatement
(oracle/ts_host.py) against scipy, closed forms and an independent
interpolation, and the library's tables (mrx_ts_snapshot_tables, through the C
ABI; no GPU needed) against the oracle's.

Reference: XRayHeatingFunctions.c:105-1511 (tables, integrals, root finding,
growth factor), ComputeTs.c:276-377 (shell plan and SFR-history weights) and
:517-716 (snapshot tables, Ly-n sums and their kink fix).  Tolerances:
* frequency integrals: QAG-61 to the reference's 1 % (its contract); measured
  against a kink-aware Gauss-Legendre reference the worst entry is 1.0e-3;
* tau_X: QAG-61 to the reference's 0.5 %;
* nu(tau_X = 1): Brent's bracket, 2 % relative, as the reference stops it;
* library vs oracle tables: 1e-12 relative (the same published algorithms,
  independently written; only rounding differs).
"""
import ctypes as C
import math
import re

import numpy as np
import pytest
from scipy import integrate, interpolate, optimize

from meraxes_amd import _lib, abi
from meraxes_amd import params as P
from meraxes_amd import synth
from oracle import oracle as O
from oracle import ts_host as H

FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
CASES = [(5, 1e-4, 2e-4), (12, 3e-3, 1e-3), (24, 2e-2, 5e-2)]


def ts_case(snap, cf, xe, N=64, L=100.0):
    z = P.snapshot_redshifts()
    fc = np.linspace(0.0, cf, len(z))
    p = P.ts_params(N, L, snap, z, fc, 8)
    p.fixed_globals = 1
    p.fixed_collapse_fraction = cf
    p.fixed_x_e_ave = xe
    return p, z, fc


@pytest.fixture(scope="module")
def tables():
    """(params, oracle snapshot, library mrx_ts_tables) for each case."""
    L = _lib.load()
    out = []
    for snap, cf, xe in CASES:
        p, z, fc = ts_case(snap, cf, xe)
        s, _, _, _ = O.ts_tables(p)
        p2, z2, fc2 = ts_case(snap, cf, xe)
        t = abi.TsTables()
        assert L.mrx_ts_snapshot_tables(C.byref(p2), cf, xe, C.byref(t)) == 0
        assert fc2[snap] == cf  # stored_fcoll[snapshot] receives the collapse fraction, as in the reference
        out.append((p, s, t, (z, fc, z2, fc2)))
    return out


# ------------------------------------------------------------------ GSL routines

def test_gk61_rule_from_its_definition():
    x = np.concatenate([H.XGK[:30], -H.XGK[:30], [0.0]])
    w = np.concatenate([H.WGK[:30], H.WGK[:30], [H.WGK[30]]])
    for k in range(0, 92):
        ex = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x**k) - ex) <= 2e-14 * max(ex, 1e-3), k
    gx, gw = np.polynomial.legendre.leggauss(30)
    pos = np.sort(gx[gx > 0])[::-1]
    np.testing.assert_allclose(H.XGK[1:30:2], pos, rtol=0, atol=1e-15)
    np.testing.assert_allclose(H.WG, gw[gx > 0][np.argsort(gx[gx > 0])[::-1]], rtol=1e-13)
    # the library's generated rule is the same rule
    src = open(_lib.HERE + "/csrc/gk61_rule.h").read()
    arr = [np.array([float(v) for v in a.split(",")]) for a in re.findall(r"\{([^}]*)\}", src)]
    np.testing.assert_allclose(arr[0], H.XGK, rtol=0, atol=1e-15)
    np.testing.assert_allclose(arr[1], H.WGK, rtol=1e-12)
    np.testing.assert_allclose(arr[2], H.WG, rtol=1e-12)


def test_qag61_matches_quad_and_the_library():
    f = lambda x: np.exp(-x) * np.sin(3 * x) / (1 + x * x)  # noqa: E731
    ref = integrate.quad(lambda x: math.exp(-x) * math.sin(3 * x) / (1 + x * x), 0, 20, epsabs=0, epsrel=1e-13)[0]
    got = H.qag61(f, 0.0, 20.0, 1e-10)[0]
    assert abs(got - ref) <= 1e-10 * abs(ref)
    g = lambda x: np.abs(x - 0.3) ** 0.5  # noqa: E731  (a kink: QAG subdivides)
    got, _, nint = H.qag61(g, 0.0, 1.0, 0.01)
    ref = (2 / 3) * (0.3**1.5 + 0.7**1.5)
    assert abs(got - ref) <= 0.01 * ref and nint > 1
    L = _lib.load()
    L.mrx_ts_qag61.restype = C.c_double
    L.mrx_ts_qag61.argtypes = [FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p]
    for fn, a, b, eps in ((f, 0.0, 20.0, 1e-10), (g, 0.0, 1.0, 0.01), (lambda x: 1 / (1e-3 + x * x), -1, 2, 1e-6)):
        cb = FN(lambda x, _: float(fn(np.array([x]))[0]))
        lib = L.mrx_ts_qag61(cb, None, a, b, eps, 1000, None)
        assert lib == pytest.approx(H.qag61(fn, a, b, eps)[0], rel=1e-13)


def test_brent_as_gsl_runs_it():
    """Brent's iteration with GSL's bracket test: the bracket reported after each
    step is the new point and the end of opposite sign, so it always holds a
    sign change; the library's root finder returns the same point."""
    L = _lib.load()
    L.mrx_ts_brent.restype = C.c_double
    L.mrx_ts_brent.argtypes = [FN, C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int]
    funcs = [(lambda x: math.cos(x) - x, 0.0, 1.0), (lambda x: x**3 - 2 * x - 5, 1.0, 3.0),
             (lambda x: math.exp(-x) - 0.03, 0.0, 40.0), (lambda x: 1.0 / (x * x * x) - 1e-6, 1.0, 1e4)]
    for f, lo, hi in funcs:
        root = optimize.brentq(f, lo, hi, xtol=1e-15, rtol=4 * np.finfo(float).eps)
        for relerr in (1e-12, 0.02):
            got, _ = H.brent_gsl(f, lo, hi, relerr)
            assert abs(got - root) <= max(relerr, 1e-14) * abs(root) * 1.0001
            assert L.mrx_ts_brent(FN(lambda x, _: f(x)), None, lo, hi, relerr, 100) == got


def test_natural_spline_matches_scipy():
    rng = np.random.default_rng(0)
    x = np.sort(rng.random(17)) * 10
    y = np.sin(x) + rng.random(17)
    sp = H.NaturalSpline(x, y)
    cs = interpolate.CubicSpline(x, y, bc_type="natural")
    for xv in np.linspace(x[0], x[-1], 97):
        assert abs(sp(xv) - cs(xv)) <= 1e-12 * (1 + abs(cs(xv)))


# ------------------------------------------------------------ tables and physics

def test_recfast_kappa_and_spectra_reproduce_their_files():
    p, _, _ = ts_case(*CASES[0])
    h = H.host(p)
    tab = np.loadtxt(P.TABLES_DIR + "/recfast_LCDM.dat")
    for z, xe, _, T in tab[::37]:
        assert h.T_RECFAST(z) == pytest.approx(float(np.float32(T)), rel=1e-12)
        assert h.xion_RECFAST(z) == pytest.approx(float(np.float32(xe)), rel=1e-12)
    for T, k in [(1.0, 1.38e-13), (30.0, 3.67e-11), (1000.0, 3.31e-10), (10000.0, 7.87e-10)]:
        assert math.exp(h.d.kappa_HH(math.log(T))) == pytest.approx(k, rel=1e-12)
    for name, n in (("kappa_eH", 20), ("kappa_pH", 17)):
        sp = getattr(h.d, name)
        for T, k in np.loadtxt(P.TABLES_DIR + f"/{name}_table.dat", max_rows=n):
            assert math.exp(sp(math.log(float(np.float32(T))))) == pytest.approx(float(np.float32(k)), rel=1e-7)
    # the normalised Pop II spectrum integrates, over each Ly-n band, to its file
    # fraction times N_ion (the normalisation of XRayHeatingFunctions.c:633-638)
    raw = np.loadtxt(P.TABLES_DIR + "/stellar_spectra.dat", max_rows=22)
    d = h.d
    for i in range(1, H.NSPEC_MAX - 1):
        a, b = d.spec_nu[i], d.spec_nu[i + 1]
        band = integrate.quad(lambda nu: d.spectral_emissivity(nu) * H.LY_ALPHA_HZ, a, b * (1 - 1e-15),
                              epsabs=0, epsrel=1e-11)[0]
        assert band == pytest.approx(float(np.float32(raw[i - 1, 1])) * p.nion_phot_per_bary, rel=2e-6)


def test_xint_interpolation_against_an_independent_bilinear():
    """interp_* of XRayHeatingFunctions.c:914-1192: float bilinear interpolation
    with clamps, on the energy bin that locate_energy_index (:1194-1207) computes
    from the table's log spacing.  That bin is the bracketing one below
    1008.88 eV; from there up the formula gives one bin less than the bracket
    (the coarse part of the table starts at index 233, the formula at 232), so
    the reference extrapolates the previous segment's line - kept for parity.
    Checked against the same lines evaluated in double."""
    p, _, _ = ts_case(*CASES[0])
    d = H.host(p).d
    E = d.E.astype(np.float64)
    rng = np.random.default_rng(1)
    En = np.concatenate([np.exp(rng.uniform(np.log(10.0), np.log(E[-1] * 0.998), 400)), [E[-1] * 2.0]])
    En = En.astype(np.float32).astype(np.float64)
    Ec = np.minimum(En, float(np.float32(0.999 * E[-1])))
    n = np.where(Ec < 1008.88, np.log(Ec / 10.0) / 1.98026273e-2, 232 + np.log(Ec / 1008.88) / 9.53101798e-2)
    n = n.astype(np.int64)
    bracket = np.searchsorted(E, Ec, side="right") - 1
    assert np.mean(n[Ec < 1008.88] == bracket[Ec < 1008.88]) > 0.99
    assert np.mean(n[Ec >= 1008.88] == bracket[Ec >= 1008.88] - 1) > 0.99
    for which in ("fheat", "nLya", "nHI", "nHeI", "nHeII"):
        tab = d.xint[which].astype(np.float64)
        for x in (1.0e-4, 3e-4, 0.02, 0.7, 0.999):
            x32, xlo, xhi = float(np.float32(x)), float(H.XHII[0]), float(H.XHII[-1])
            xc = float(np.float32(xhi * 0.999)) if x32 > xhi * 0.999 else (
                float(np.float32(1.001 * xlo)) if x32 < xlo else x32)
            m = np.searchsorted(H.XHII.astype(np.float64), xc, side="right") - 1
            line = lambda row: tab[row, n] + (tab[row, n + 1] - tab[row, n]) / (E[n + 1] - E[n]) * (Ec - E[n])  # noqa
            lo_row, hi_row = line(m), line(m + 1)
            x0, x1 = float(H.XHII[m]), float(H.XHII[m + 1])
            want = lo_row + (hi_row - lo_row) * (xc - x0) / (x1 - x0)
            got = d.interp(which, En.astype(np.float32), np.float32(x)).astype(np.float64)
            np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6 * max(1.0, np.abs(tab).max()),
                                       err_msg=f"{which} x_e={x}")
    # below the first tabulated energy: f_heat = 1, the others 0
    assert d.interp("fheat", np.array([5.0], np.float32), np.float32(0.5))[0] == 1.0
    assert d.interp("nHI", np.array([5.0], np.float32), np.float32(0.5))[0] == 0.0


def test_cosmology_closed_forms():
    p, _, _ = ts_case(*CASES[0])
    h = H.host(p)
    H0 = H.HUBBLE * P.HUBBLE_H
    Om, Ol = P.OMEGA_M, P.OMEGA_L
    for z in (0.0, 6.0, 9.5, 20.0):
        opz = float(np.float32(1) + np.float32(z))  # dtdz and hubble take a float z, 1 + z in float
        Hz = H0 * math.sqrt(Om * opz**3 + Ol)
        assert h.hubble(z) == pytest.approx(Hz, rel=1e-13)
        assert h.dtdz(z) == pytest.approx(-1 / (opz * Hz), rel=1e-9)  # flat LCDM: dt/dz = -1/((1+z)H)
    rhoc = 3 * H0**2 / (8 * math.pi * H.GRAVITY)
    assert h.No == pytest.approx(rhoc * Om * P.BARYON_FRAC * (1 - P.Y_HE) / H.PROTONMASS, rel=1e-14)


def test_growth_factor():
    """dicke (XRayHeatingFunctions.c:1469-1504) is the Liddle et al. fit: 1 today,
    1/(1+z) for Omega_m = 1, within 1 % of the exact linear growth integral; and
    ddicke_dz is its forward difference with the float step 1e-10, within that
    difference's cancellation error of the fit's exact derivative."""
    import mpmath as mp

    p, _, _ = ts_case(*CASES[0])
    h = H.host(p)
    assert h.dicke(0.0) == 1.0
    Om, Ol = P.OMEGA_M, P.OMEGA_L

    def D_exact(a):
        E = lambda a_: math.sqrt(Om / a_**3 + Ol)  # noqa: E731
        return E(a) * integrate.quad(lambda a_: 1 / (a_ * E(a_)) ** 3, 0, a, epsabs=0, epsrel=1e-12)[0]

    for z in (6.0, 10.0, 20.0):
        assert h.dicke(z) == pytest.approx(D_exact(1 / (1 + z)) / D_exact(1.0), rel=0.01)
        fit = lambda zz: (lambda om: 2.5 * om / (1 / 70.0 + om * (209 - om) / 140.0 + om ** (mp.mpf(4) / 7)))(  # noqa
            Om * (1 + zz) ** 3 / (Ol + Om * (1 + zz) ** 3)) / (1 + zz)
        d0 = fit(mp.mpf(0))
        assert h.dicke(z) == pytest.approx(float(fit(mp.mpf(z)) / d0), rel=1e-14)
        exact_der = float(mp.diff(lambda zz: fit(zz) / d0, mp.mpf(z)))
        # forward difference at dz = (float)1e-10: truncation ~dz|D''| plus cancellation of a few ulp of D
        dz = float(np.float32(1e-10))
        bound = 8 * np.finfo(float).eps * h.dicke(z) / dz + dz * abs(exact_der)
        assert abs(h.ddicke_dz(z) - exact_der) <= bound
    p1, _, _ = ts_case(*CASES[0])
    p1.omega_m, p1.omega_lambda = 1.0, 0.0
    assert H.TsHost(p1, h.d).dicke(7.0) == 1.0 / 8.0


def test_tauX_against_quad():
    """tau_X (QAG-61 at 0.5 %) against scipy's quad of the optical-depth integrand
    written out here: c dt/dz n_b (1+z)^3 x_HI-filling sigma(nu (1+z)/(1+z_p))."""
    snap, cf, xe = CASES[1]
    p, z, fc = ts_case(snap, cf, xe)
    h = H.host(p)
    h.plan()
    fcs = fc.copy()
    fcs[snap] = cf
    zp = z[snap]
    for zpp, nu_ev in ((zp + 0.3, 300.0), (zp + 2.0, 600.0), (zp + 6.0, 1200.0)):
        nu = nu_ev * H.NU_OVER_EV
        got = h.tauX(nu, xe, xe, zp, zpp, fcs)

        def integrand(zh):
            fcoll = max(fcs[snap - 1] + (zh - z[snap - 1]) * (fcs[snap] - fcs[snap - 1]) / (z[snap] - z[snap - 1]),
                        0.0)
            ff = 1.0 if fcoll < 1e-20 else max(1 - p.reion_efficiency * fcoll / (1 - xe), 1e-4)
            nuh = nu / (1 + zp) * (1 + zh)
            sig = (h.f_H * (1 - xe) * H.HI_ion_crosssec(nuh) + h.f_He * (1 - xe) * H.HeI_ion_crosssec(nuh) +
                   h.f_He * xe * H.HeII_ion_crosssec(nuh))
            return H.SPEED_OF_LIGHT * h.dtdz(zh) * h.N_b0 * (1 + zh) ** 3 * ff * float(sig)

        ref = integrate.quad(integrand, zpp, zp, epsabs=0, epsrel=1e-10, limit=200)[0]
        assert got == pytest.approx(ref, rel=5e-3)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_nu_tau_one_within_brents_bracket(tables, case):
    """The lower frequency limit is where tau_X = 1, to Brent's 2 % bracket
    (nu_tau_one, XRayHeatingFunctions.c:238-298), or the X-ray threshold."""
    p, s, t, (z, fc, _, _) = tables[case]
    h = H.host(p)
    h.plan()
    fcs = fc.copy()
    fcs[p.snapshot] = s.collapse_fraction
    thr = p.nu_xray_gal_threshold * H.NU_OVER_EV
    for ct in range(0, s.nsteps, 3):
        prev = s.zp if ct == 0 else s.zpp_edge[ct - 1]
        zpp = (s.zpp_edge[ct] + prev) * 0.5
        f = lambda nu: h.tauX(nu, s.x_e_ave, s.x_e_ave, s.zp, zpp, fcs) - 1  # noqa: E731
        if f(H.HEI_NUIONIZATION) < 0:
            root = H.HEI_NUIONIZATION
        else:
            root = optimize.brentq(f, H.HEI_NUIONIZATION, 1e6 * H.NU_OVER_EV, rtol=1e-12)
        nu1 = h.nu_tau_one(s.zp, zpp, s.x_e_ave, s.x_e_ave, fcs)
        assert abs(nu1 - root) <= 0.02 * root
        assert s.nu_lower[ct] == max(nu1, thr)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_frequency_integrals_against_kink_aware_quadrature(tables, case):
    """Every freq_int_{heat,ion,lya} entry (QAG-61 at 1 %, XRayHeatingFunctions.c:755-800)
    against a Gauss-Legendre sum split at every tabulated energy of every species
    (the integrand is smooth between them)."""
    p, s, _, _ = tables[case]
    h = H.host(p)
    E = h.d.E.astype(np.float64)
    gx, gw = np.polynomial.legendre.leggauss(24)
    hi = p.nu_xray_max * H.NU_OVER_EV
    worst = 0.0
    for ct in range(0, s.nsteps, 4):
        lo = s.nu_lower[ct]
        pts = {lo, hi}
        for th in (H.NUIONIZATION, H.HEI_NUIONIZATION, H.HEII_NUIONIZATION):
            for nu in np.concatenate([E, [0.999 * E[-1]]]) * H.NU_OVER_EV + th:
                if lo < nu < hi:
                    pts.add(float(nu))
        pts = np.array(sorted(pts))
        a, b = pts[:-1], pts[1:]
        xs = ((0.5 * (a + b))[:, None] + (0.5 * (b - a))[:, None] * gx[None, :]).ravel()
        wts = (gw[None, :] * (0.5 * (b - a))[:, None]).ravel()
        for x in range(0, H.NXHII, 3):
            for flag, tab in ((0, s.freq_heat), (1, s.freq_ion), (2, s.freq_lya)):
                ref = float(np.sum(h.nu_integrand(flag, float(H.XHII[x]))(xs) * wts))
                if flag == 2:
                    ref = ref * H.SPEED_OF_LIGHT / (4 * math.pi) / H.LY_ALPHA_HZ / h.hubble(s.zp)
                err = abs(tab[x, ct] - ref) / abs(ref)
                worst = max(worst, err)
                assert err <= 0.01, (ct, x, flag)
    print(f"case {case}: worst frequency-integral error {worst:.2e}")


def _emissivity_from_file(nion):
    """The stellar spectrum of the file, normalised by numerical band integrals."""
    raw = np.loadtxt(P.TABLES_DIR + "/stellar_spectra.dat", max_rows=22)
    n = raw[:, 0]
    nu = (4.0 / 3.0 * (1 - 1 / n**2)).astype(np.float32).astype(np.float64)
    N0 = raw[:, 1].astype(np.float32).astype(np.float64)
    al = raw[:, 2].astype(np.float32).astype(np.float64)
    norm = N0.copy()
    for i in range(len(n) - 1):
        band = integrate.quad(lambda v: v ** al[i], nu[i], nu[i + 1], epsrel=1e-13)[0]
        norm[i] = N0[i] * nion / band

    def eps(v):
        i = np.searchsorted(nu, v, side="right") - 1
        i = len(n) - 1 if (i < 0 or i >= len(n) - 1) else i
        return norm[i] * v ** al[i] / H.LY_ALPHA_HZ

    return eps


@pytest.mark.parametrize("case", range(len(CASES)))
def test_sum_lyn_and_kink_fix(tables, case):
    """sum_lyn (ComputeTs.c:586-623) against a direct sum of f_recycle times the
    band-normalised spectrum over the Ly-n lines seen from z'' (those with
    z'' <= z_max(n)); and the kink fix (:625-663) against its closed form: the
    fraction of the shell pair's redshift span still below z_max(2)."""
    p, s, _, _ = tables[case]
    eps = _emissivity_from_file(p.nion_phot_per_bary)
    zp = s.zp
    zpf = float(np.float32(zp))
    kink = getattr(s, "kink_shell", None)
    for ct in range(s.nsteps):
        prev = zp if ct == 0 else s.zpp_edge[ct - 1]
        zpp = (s.zpp_edge[ct] + prev) * 0.5
        want = 0.0
        for n in range(23, 1, -1):
            zmax = (1 + zpf) * (1 - (n + 1) ** -2) / (1 - n**-2) - 1
            if zpp <= zmax * (1 + 1e-7):
                want += H.frecycle(n) * eps((1 - n**-2) / 0.75 * (1 + zpp) / (1 + zp))
        if ct == kink:
            continue
        assert s.sum_lyn[ct] == pytest.approx(want, rel=3e-6, abs=1e-300)
    if kink is not None:
        zmax2 = (1 + zpf) * (1 - 1 / 9) / (1 - 1 / 4) - 1
        h = H.host(p)
        prev_zpp = s.zpp_edge[kink - 1]
        tmax = (prev_zpp - (s.R[kink] - s.R[kink - 1]) * H.MPC / h.drdz(prev_zpp) + prev_zpp) * 0.5
        e2 = s.zpp_edge[kink - 2]
        tmin = (e2 - (s.R[kink - 1] - s.R[kink - 2]) * H.MPC / h.drdz(e2) + e2) * 0.5
        frac = (zmax2 - tmin) / (tmax - tmin)
        assert abs(s.kink_weight - frac) <= 1.5e-3
        assert s.sum_lyn[kink] == s.kink_weight * s.sum_lyn[kink - 1]


def _slot_coefficients(ops, nsteps):
    """Each shell's working grid as a combination of SFR-history slots (slot 0 is
    the current grid), separately for the first half of the padded array (which
    the zero/scale operations touch) and the rest."""
    lo, hi = {0: 1.0}, {0: 1.0}
    out = []
    k = 0
    for ct in range(nsteps):
        while k < len(ops) and ops[k][0] == ct:
            _, kind, slot, w = ops[k]
            if kind == 0:
                lo = {}
            elif kind == 1:
                lo = {j: v * w for j, v in lo.items()}
            else:
                wf = float(np.float32(w))
                if kind == 2:
                    lo, hi = {slot: wf}, {slot: wf}
                else:
                    lo = dict(lo)
                    hi = dict(hi)
                    lo[slot] = lo.get(slot, 0.0) + wf
                    hi[slot] = hi.get(slot, 0.0) + wf
            k += 1
        out.append((dict(lo), dict(hi)))
    return out


@pytest.mark.parametrize("case", range(len(CASES)))
def test_sfr_history_weights(tables, case):
    """The SFR-history operations (ComputeTs.c:296-376) make each shell the
    dt/dz-weighted mix of the snapshots whose intervals overlap it: the SFR at
    redshift z in (ZZ[s-j], ZZ[s-j-1]] comes from slot j.  The reference weighs
    each piece by the midpoint rule (the first shell's current-snapshot piece
    by dt/dz at z_p, an end-point rule) over the shell's own midpoint rule, so
    each weight is checked against the exact integrals of |dt/dz| to those
    rules' error bounds.  Beyond snapshot 0 there is no source: the reference
    zeroes the first half of the working grid there (:303-304), which is the
    part checked (the zero and scale operations never touch the rest)."""
    p, s, _, (z, _, _, _) = tables[case]
    snap = p.snapshot
    ZZ = z[:snap + 1]
    H0 = H.HUBBLE * p.hubble_h
    f = lambda zz: 1.0 / ((1 + zz) * H0 * math.sqrt(p.omega_m * (1 + zz) ** 3 + p.omega_lambda))  # noqa: E731
    df = lambda zz: (f(zz + 1e-4) - f(zz - 1e-4)) / 2e-4  # noqa: E731
    d2f = lambda zz: (f(zz + 1e-3) - 2 * f(zz) + f(zz - 1e-3)) / 1e-6  # noqa: E731
    quad = lambda a, b: integrate.quad(f, a, b, epsrel=1e-13)[0]  # noqa: E731

    def mid_bound(a, b):  # |midpoint - exact| <= (b - a)^3 / 24 max|f''|
        return (b - a) ** 3 / 24 * max(abs(d2f(t)) for t in np.linspace(a, b, 9))

    coeffs = _slot_coefficients(s.ops, s.nsteps)
    for ct in range(s.nsteps):
        zlo = s.zp if ct == 0 else s.zpp_edge[ct - 1]
        zhi = s.zpp_edge[ct]
        total, total_err = quad(zlo, zhi), mid_bound(zlo, zhi)
        want, err = {}, {}
        for j in range(snap):
            a, b = max(zlo, ZZ[snap - j]), min(zhi, ZZ[snap - j - 1])
            if b > a:
                want[j] = quad(a, b) / total
                piece_err = mid_bound(a, b)
                if ct == 0 and j == 0:  # end-point rule: |f(a)(b - a) - exact| <= (b - a)^2 / 2 max|f'|
                    piece_err = (b - a) ** 2 / 2 * max(abs(df(t)) for t in np.linspace(a, b, 9))
                err[j] = (piece_err + want[j] * total_err) / total * 1.05 + 1e-6
        got = coeffs[ct][0]
        assert set(k for k, v in got.items() if v != 0) <= set(want), (ct, got, want)
        for j, w in want.items():
            assert abs(got.get(j, 0.0) - w) <= err[j], (ct, j, got.get(j, 0.0), w, err[j])


@pytest.mark.parametrize("case", range(len(CASES)))
def test_library_tables_equal_the_oracles(tables, case):
    p, s, t, _ = tables[case]
    ns = s.nsteps
    assert t.nsteps == ns and t.homogeneous == 0 and t.no_light == s.no_light
    for name in ("R", "zpp_edge", "zpp", "dzpp", "dt_dzpp", "zpp_pow", "sum_lyn"):
        np.testing.assert_allclose(np.array(getattr(t, name)[:ns]), np.array(getattr(s, name)), rtol=1e-15,
                                   atol=0, err_msg=name)
    np.testing.assert_allclose(np.array(t.nu_lower[:ns]), np.array(s.nu_lower), rtol=1e-12, atol=0)
    for name in ("freq_heat", "freq_ion", "freq_lya"):
        got = np.array([[getattr(t, name)[x][c] for c in range(ns)] for x in range(14)])
        np.testing.assert_allclose(got, getattr(s, name), rtol=1e-12, atol=0, err_msg=name)
    for a, b in (("const_zp_prefactor", "const_zp_prefactor"), ("growth_factor_zp", "growth_factor_zp"),
                 ("dgrowth_factor_dzp", "dgrowth_factor_dzp"), ("dt_dzp", "dt_dzp"), ("zp", "zp"), ("dzp", "dzp")):
        assert getattr(t, a) == pytest.approx(getattr(s, b), rel=1e-15), a
    ops = [(t.op_shell[k], t.op_kind[k], t.op_slot[k], t.op_weight[k]) for k in range(t.n_ops)]
    assert len(ops) == len(s.ops)
    for (c1, k1, sl1, w1), (c2, k2, sl2, w2) in zip(ops, s.ops):
        assert (c1, k1, sl1) == (c2, k2, sl2) and w1 == pytest.approx(w2, rel=1e-15)


def test_plan_limits_shell_operations():
    """A shell may span many closely spaced snapshots: the library plans up to
    256 operations per shell (the device applies them 24 per launch) and
    refuses more (MRX_EUNSUPPORTED) instead of dropping any."""
    L = _lib.load()
    z = np.linspace(20.0, 8.0, 1000)
    fc = np.linspace(0, 1e-3, 1000)
    p = P.ts_params(32, 100.0, 700, z, fc, 256)
    p.fixed_globals, p.fixed_collapse_fraction, p.fixed_x_e_ave = 1, 1e-3, 1e-3
    t = abi.TsTables()
    assert L.mrx_ts_snapshot_tables(C.byref(p), 1e-3, 1e-3, C.byref(t)) == 0
    per_shell = np.bincount([t.op_shell[k] for k in range(t.n_ops)], minlength=t.nsteps)
    assert per_shell.max() > 24  # more than one launch's worth in a shell
    s = H.host(p).plan()
    assert [(c, k, sl) for c, k, sl, _ in s.ops] == [(t.op_shell[i], t.op_kind[i], t.op_slot[i])
                                                      for i in range(t.n_ops)]
    # eight shells over 6000 snapshots: several hundred operations in the outer shells
    zf = np.linspace(20.0, 8.0, 6000)
    fcf = np.linspace(0, 1e-3, 6000)
    pf = P.ts_params(32, 100.0, 4000, zf, fcf, 8192, num_filter_steps=8)
    sf = H.host(pf).plan()
    assert np.bincount([o[0] for o in sf.ops]).max() > 256
    assert L.mrx_ts_snapshot_tables(C.byref(pf), 1e-3, 1e-3, C.byref(t)) == -4


# ------------------------------------------------------ per-cell closed forms

def test_no_light_evolution_closed_form():
    """With no sources the x_e / T_K step is recombination, Compton and adiabatic
    cooling only (evolveInt with NO_LIGHT), and T_S the collisions-only closed form."""
    N, L_ = 16, 25.0
    z = P.snapshot_redshifts()
    fc = np.zeros(len(z))
    zero = synth.to_padded(np.zeros((N, N, N), np.float32))
    hist = np.zeros((4, N, N, 2 * (N // 2 + 1)), np.float32)
    xe0 = np.full((N, N, 2 * (N // 2 + 1)), 3e-4, np.float32)
    T0 = np.full((N, N, N), 12.0, np.float32)
    p = P.ts_params(N, L_, 2, z, fc, 4)
    o = O.compute_ts(p, zero, zero, zero, hist, xe0, T0)
    assert o["no_light"] == 1 and o["collapse_fraction"] == 0.0
    h = H.host(p)
    zp = float(np.float32(z[2]))
    dzp = z[2] - z[1]
    dt_dzp = h.dtdz(z[2])
    xe, T = 3e-4, 12.0
    n_b = h.N_b0 * (1 + zp) ** 3
    lt = math.log(T / 1.1604505e4)  # case-A recombination, Abel et al. (1997) fit
    cs = [-28.6130338, -0.72411256, -2.02604473e-2, -2.38086188e-3, -3.21260521e-4, -1.42150291e-5,
          4.98910892e-6, 5.75561414e-7, -1.85676704e-8, -3.07113524e-9]
    alphaA = math.exp(sum(c * lt**k for k, c in enumerate(cs)))
    assert 3.5e-13 < math.exp(sum(c * math.log(1.0) ** k for k, c in enumerate(cs))) < 5e-13
    dxe = dt_dzp * (-alphaA * 2.0 * xe * xe * h.f_H * n_b)
    Trad = 2.728 * (1 + zp)
    Hz = h.hubble(zp)
    dcomp = -1.51e-4 * xe / (1 + xe + h.f_He) / (Hz / (H.HUBBLE * P.HUBBLE_H)) / P.HUBBLE_H * Trad**4 / (1 + zp) * (
        Trad - T)
    dT = dcomp + (-dxe * T / (1 + xe)) + 2.0 / (1 + zp) * T
    assert o["x_e_box_prev"][0, 0, 0] == pytest.approx(xe + dxe * dzp, rel=1e-6)
    assert o["Tk_box"][0, 0, 0] == pytest.approx(T + dT * dzp, rel=1e-6)
    TK = float(o["Tk_box"][0, 0, 0])
    xen = float(o["x_e_box_prev"][0, 0, 0])
    k = lambda sp, t_: math.exp(sp(math.log(t_)))  # noqa: E731  (inside every table's range here)
    nz3 = (1 + zp) ** 3
    xc = 0.0628 / Trad * ((1 - xen) * h.No * nz3 * k(h.d.kappa_HH, TK) + xen * h.N_b0 * nz3 * k(h.d.kappa_eH, TK) +
                          xen * h.No * nz3 * k(h.d.kappa_pH, TK)) / 2.85e-15
    TS = (1 + xc) / (1 / Trad + xc / TK)
    assert o["TS_box"][0, 0, 0] == pytest.approx(TS, rel=1e-6)


def test_shell_ladder(tables):
    p, s, _, (z, _, _, _) = tables[0]
    box = p.box_size / p.hubble_h
    R = np.array(s.R)
    assert R[0] == pytest.approx(0.620350491 * box / 64, rel=1e-12)
    assert R[-1] * (500 / R[0]) ** (1 / 40) == pytest.approx(500, rel=1e-5)  # R_XLy_MAX after 40 steps
    assert np.all(np.diff(s.zpp_edge) > 0)  # light-cone edges recede in redshift
    assert s.zpp_edge[0] > s.zp
