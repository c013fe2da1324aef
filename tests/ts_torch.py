"""Float64 torch restatement of one snapshot's differentiable map, for tests.

Given the snapshot constants from the oracle (`oracle.ts_export`), it rebuilds
in torch float64, with autograd: the SFR working-grid operations and the 40
spin-temperature shells (sharp-k windows evaluated in float as filter() does,
reionization.c:2110-2170), the per-cell Euler step of evolveInt and get_Ts
(ComputeTs.c:743-900, XRayHeatingFunctions.c:1241-1558), the partially ionised
xH on the unfiltered last step with the ionisation mask held fixed
(find_HII_bubbles.c:317-443), delta_T (BrightnessTemperature.c:49-124) and the
binned power of test_gpu_adjoint.  It shares no code with the HIP adjoint.
"""
import numpy as np
import torch

TCMB = 2.728
MSUN = 1.989e33


def sharpk_mask(N, box_size, R):
    """filter()'s sharp-k window in its float arithmetic, on the rfft grid."""
    f32 = np.float32
    dk = f32(2.0 * np.pi / float(f32(box_size)))
    n = np.arange(N)
    kx = (np.where(n > N // 2, n - N, n).astype(f32) * dk).astype(f32)
    kz = (np.arange(N // 2 + 1).astype(f32) * dk).astype(f32)
    k2 = (kx[:, None, None] * kx[:, None, None] + kx[None, :, None] * kx[None, :, None]).astype(f32)
    k2 = (k2 + kz[None, None, :] * kz[None, None, :]).astype(f32)
    kmag = np.sqrt(k2).astype(f32)
    kR = (kmag * f32(R)).astype(f32)
    kR = (kR.astype(np.float64) * 0.413566994).astype(f32)
    return (kR <= 1.0).astype(np.float64)


def shells(c, sfr, hist, N, box_size, dev):
    """The clamped shells [ns, N, N, N] from the padded sfr and history grids."""
    W = sfr[:, :, :N]
    ops = {}
    for ct, kind, slot, w in c["ops"]:
        ops.setdefault(ct, []).append((kind, slot, w))
    out = []
    for ct in range(c["nsteps"]):
        for kind, slot, w in ops.get(ct, []):
            if kind == 0:
                W = torch.cat([torch.zeros_like(W[: N // 2]), W[N // 2:]])
            elif kind == 1:
                W = torch.cat([W[: N // 2] * w, W[N // 2:]])
            else:
                h = hist[slot].reshape(-1)[: N**3].reshape(N, N, N) * w  # read at the unpadded index
                W = h if kind == 2 else W + h
        if ct == 0:
            sh = W
        else:
            m = torch.as_tensor(sharpk_mask(N, box_size, float(np.float32(c["R"][ct]))), device=dev)
            sh = torch.fft.irfftn(torch.fft.rfftn(W) * m, s=(N, N, N))
        out.append(torch.where(sh > 0, sh, torch.zeros_like(sh)))  # max(v, 0), derivative 0 at v = 0
    return torch.stack(out)


def _spline(sp, x):
    xs = torch.as_tensor(sp["x"], device=x.device)
    n = len(sp["x"])
    lo = torch.clamp(torch.searchsorted(xs, x.detach(), right=True) - 1, 0, n - 2)
    dx = x - xs[lo]
    g = lambda k: torch.as_tensor(sp[k], device=x.device)[lo]  # noqa: E731
    return g("y") + dx * (g("b") + dx * (g("c") + dx * g("d")))


def _kappa_HH(sp, TK):
    lt = torch.log(TK)
    x0, xn, y0, yn = sp["x"][0], sp["x"][-1], sp["y"][0], sp["y"][-1]
    ext = np.log(np.exp(yn)) + 0.381 * (lt - xn)  # log(exp(y_n) (T / exp(x_n))^0.381)
    ans = torch.where(lt < x0, torch.full_like(lt, y0), torch.where(lt > xn, ext, _spline(sp, lt)))
    return torch.exp(ans)


def _kappa_file(sp, T):
    lt = torch.log(T)
    x, y = sp["x"], sp["y"]
    ext = y[-1] + (y[-1] - y[-2]) / (x[-1] - x[-2]) * (lt - x[-1])
    ans = torch.where(lt < x[0], torch.full_like(lt, y[0]), torch.where(lt > x[-1], ext, _spline(sp, lt)))
    return torch.exp(ans)


def _alpha_A(T):
    lt = torch.log(T / 1.1604505e4)
    cs = [-28.6130338, -0.72411256, -2.02604473e-2, -2.38086188e-3, -3.21260521e-4, -1.42150291e-5,
          4.98910892e-6, 5.75561414e-7, -1.85676704e-8, -3.07113524e-9]
    acc = torch.full_like(lt, cs[0])
    for k in range(1, 10):
        acc = acc + cs[k] * lt**k
    return torch.exp(acc)


def get_Ts(c, dl, TK, xe, Ja):
    zp = float(np.float32(c["zp"]))
    Trad = TCMB * (1.0 + zp)
    P3 = (1.0 + zp) ** 3
    T21_Trad = 0.0628 / Trad
    nH = (1.0 - xe) * c["No"] * P3 * (1.0 + dl)
    xHI = T21_Trad * nH * _kappa_HH(c["kHH"], TK) / 2.85e-15
    ne = xe * c["N_b0"] * P3 * (1.0 + dl)
    xel = T21_Trad * ne * _kappa_file(c["keH"], TK) / 2.85e-15
    npr = xe * c["No"] * P3 * (1.0 + dl)
    xpr = T21_Trad * npr * _kappa_file(c["kpH"], TK) / 2.85e-15
    xc = xHI + xel + xpr
    TGP = 1.342881e-7 / c["H"] * c["No"] * (1.0 + zp) ** 3
    XA = 1.66e11 / (1.0 + zp)
    TS = torch.full_like(TK, Trad)
    TSold = torch.zeros_like(TK)
    lya = (Ja > 1e-20).detach()
    tgp = TGP * (1.0 + dl) * (1.0 - xe)
    for _ in range(200):
        run = lya & ((TS - TSold).abs() / TS > 1e-3).detach()
        if not bool(run.any()):
            break
        xi = (1.0e-7 * tgp / TK / TK) ** (1.0 / 3.0)
        S = 1.0 - 0.0631789 / TK + 0.115995 / TK / TK - 0.401403 / TS / TK + 0.336463 / TS / TK / TK
        S = S / (1.0 + 2.98394 * xi + 1.53583 * xi * xi + 3.85289 * xi * xi * xi)
        xa = XA * S * Ja
        ce = 1.0 / TK + 0.405535 / TK * (1.0 / TS - 1.0 / TK)
        TSinv = (1.0 / Trad + xa * ce + xc / TK) / (1.0 + xa + xc)
        TSold = torch.where(run, TS, TSold)
        TS = torch.where(run, 1.0 / TSinv, TS)
    closed = 1.0 / ((1.0 / Trad + xc / TK) / (1.0 + xc))
    TS = torch.where(lya, TS, closed)
    return TS.abs()


def ts_step(c, dl, xe0, T, sh, pixel_volume, MT, L3):
    """x_e, T_K, T_S after the step (float64), sh: the clamped shells [ns, N, N, N]."""
    XH = c["XH"]
    hi_c = float(XH[13]) * 0.999
    lo_v = 1.001 * float(XH[0])
    x_call = torch.where(xe0 > hi_c, torch.full_like(xe0, hi_c), torch.where(xe0 < float(XH[0]),
                                                                             torch.full_like(xe0, lo_v), xe0))
    xf = x_call.detach().float().cpu().numpy()
    # the forward's `lo = 13; while (x < XH[lo]) lo--` on float values
    lo = np.clip(np.searchsorted(XH, xf, side="right") - 1, 0, 12).astype(np.int64)
    lo_t = torch.as_tensor(lo, device=xe0.device)
    XH64 = torch.as_tensor(XH.astype(np.float64), device=xe0.device)
    dxh = torch.as_tensor((XH[1:] - XH[:-1]).astype(np.float64), device=xe0.device)  # float differences
    fr = x_call - XH64[lo_t]
    ns = c["nsteps"]
    SFR = sh / pixel_volume * MT * L3 / MSUN  # [ns, ...]

    def interp(name):
        tab = torch.as_tensor(c[name], device=xe0.device)  # [14, ns]
        slope = (tab[1:] - tab[:-1]) / dxh[:, None]  # [13, ns]
        return (slope[lo_t] * fr[..., None] + tab[lo_t]).permute(3, 0, 1, 2)  # [ns, ...]

    w = torch.as_tensor(c["dt_dzpp"] * c["dzpp"] * c["zpp_pow"], device=xe0.device)[:, None, None, None]
    zp = float(np.float32(c["zp"]))
    if c["no_light"]:
        H = I = Ly = St = torch.zeros_like(xe0)
    else:
        H = (w * SFR * interp("freq_heat")).sum(0)
        I = (w * SFR * interp("freq_ion")).sum(0)
        Ly = (w * SFR * interp("freq_lya")).sum(0)
        wl = torch.as_tensor((1.0 + zp) ** 2 * (1.0 + c["zpp"]) * c["sum_lyn"] * c["dt_dzpp"] * c["dzpp"],
                             device=xe0.device)[:, None, None, None]
        St = (SFR * wl).sum(0)
    cpf = c["const_zp_prefactor"]
    n_b = c["N_b0"] * (1.0 + zp) ** 3 * (1.0 + dl)
    dxheat = H * cpf
    dxion = I * cpf
    dxlya = Ly * cpf * n_b
    dstarlya = St * (2.9979e10 / (4.0 * np.pi)) / (1.6726e-24 / MSUN)
    x_e = xe0
    sink = _alpha_A(T) * 2.0 * x_e * x_e * c["f_H"] * n_b
    dxe_dzp = c["dt_dzp"] * (dxion - sink)
    dadia = 3.0 / (1.0 + zp) + torch.where(dl.abs() > 1e-7, c["dgrowth"] / (1.0 / dl + c["growth"]),
                                            torch.zeros_like(dl))
    dadia = dadia * (2.0 / 3.0) * T
    dspec = -dxe_dzp * T / (1.0 + x_e)
    Trad = TCMB * (1.0 + zp)
    HH = c["H"] / (3.2407789e-18 * c["hubble_h"])
    dcomp = (-1.51e-4) * (x_e / (1.0 + x_e + c["f_He"])) / HH / c["hubble_h"] * Trad**4 / (1.0 + zp) * (Trad - T)
    dxheat_dzp = dxheat * c["dt_dzp"] * 2.0 / 3.0 / 1.3806e-16 / (1.0 + x_e)
    der1 = dxheat_dzp + dcomp + dspec + dadia
    der2 = dxlya + dstarlya
    xn = x_e + dxe_dzp * c["dzp"]
    xn = torch.where(xn > 1, torch.full_like(xn, 1 - 1e-7), torch.where(xn < 0, torch.zeros_like(xn), xn))
    Tn = torch.where(T < 5e4, T + der1 * c["dzp"], T)
    Tn = torch.where(Tn < 0, torch.full_like(Tn, TCMB * (1 + c["zp"])), Tn)
    TS = get_Ts(c, dl, Tn, xn, der2)
    return xn, Tn, TS
