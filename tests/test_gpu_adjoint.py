"""GPU: the power spectrum and the hand-written adjoint of one snapshot.

* Compute_PS on the GPU against a float64 numpy restatement (bins, counts, power);
* the power-spectrum adjoint against torch float64 autograd of the same binned
  power, every cell;
* the snapshot adjoint without spin temperature against float64 autograd of
  delta_T = C xH (1 + delta) with the GPU forward's ionisation mask held fixed,
  through a power-spectrum loss, every cell;
* with spin temperature: central finite differences (eps = 1e-3) of the GPU
  forward's loss on 64 random cells whose threshold margin is at least 1e-3,
  with the shell tables' grid means fixed in both runs; and every gradient
  (deltax, stars, sfr, SFR history, x_e_prev, T_K) against float64 autograd of
  an independent torch restatement (tests/ts_torch.py) at 64^3.
"""
import numpy as np
import pytest

from meraxes_amd import autograd as MA
from meraxes_amd import params as P
from meraxes_amd import reion, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def ps_bins(N, box_mpc, k_factor):
    """Bin index of every half-spectrum mode, Compute_PS's float arithmetic."""
    dk = np.float32(2.0 * np.pi / box_mpc)
    kmax = np.float32(dk * np.float32(N))
    ceils = []
    c = dk
    while c < kmax:
        ceils.append(c)
        c = np.float32(c * np.float32(k_factor))
    n = np.arange(N)
    nw = np.where(n > N // 2, n - N, n).astype(np.float32)
    kx = (nw * dk).astype(np.float32)
    kz = (np.arange(N // 2 + 1).astype(np.float32) * dk).astype(np.float32)
    k2 = (kx[:, None, None] * kx[:, None, None] + kx[None, :, None] * kx[None, :, None]).astype(np.float32)
    k2 = (k2 + kz[None, None, :] * kz[None, None, :]).astype(np.float32)
    kmag = np.sqrt(k2.astype(np.float64)).astype(np.float32)
    b = np.full(kmag.shape, -1, np.int64)
    fl = np.float32(0.0)
    for i, ce in enumerate(ceils):
        b[(kmag >= fl) & (kmag < ce) & (b < 0)] = i
        fl = ce
    return b, kmag.astype(np.float64), len(ceils)


def ps_torch(dT, box_mpc, k_factor, bins, kmag, nb):
    """Differentiable float64 binned power (the math of Compute_PS)."""
    N = dT.shape[0]
    V = float(np.float32(box_mpc) ** 3)
    D = (dT - dT.mean()) * (V / N**3)
    F = torch.fft.rfftn(D)
    w = torch.as_tensor(kmag**3 / (2 * np.pi**2 * V), dtype=torch.float64, device=dT.device)
    amp2 = (F.real**2 + F.imag**2) * w
    bt = torch.as_tensor(bins, device=dT.device)
    out = []
    for b in range(nb):
        m = bt == b
        out.append(amp2[m].sum() / m.sum())
    return torch.stack(out)


def tb_const(z):
    """const_factor of BrightnessTemperature.c:46-52 (float inputs, as the reference)."""
    f32 = np.float32
    zf, h, Om = f32(z), f32(P.HUBBLE_H), f32(P.OMEGA_M)
    Ob = f32(float(Om) * P.BARYON_FRAC)
    a = float(f32(f32(Ob * h) * h)) / 0.023
    return float(f32(27.0 * a * np.sqrt((0.15 / float(Om) / float(h) / float(h)) * (1.0 + float(zf)) / 10.0)))


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("N", [32, 64])
def test_power_spectrum_matches_float64(ctx, N):
    box = 100.0 / P.HUBBLE_H
    d, s, w = synth.snapshot_inputs(N, 5)
    dT = torch.from_numpy(10.0 + 5.0 * synth.from_padded(d)).cuda()
    g = reion.power_spectrum(ctx, dT, box)
    bins, kmag, nb = ps_bins(N, box, 1.35)
    ref = ps_torch(dT.double(), box, 1.35, bins, kmag, nb).cpu().numpy()
    counts = np.array([(bins == b).sum() for b in range(nb)])
    print(f"N={N}: {nb} bins, power rel L2 {rel_l2(g['power'], ref):.2e}", flush=True)
    assert len(g["power"]) == nb
    assert np.array_equal(g["counts"], counts)
    assert rel_l2(g["power"], ref) < 1e-5


@pytest.mark.parametrize("N", [32, 64])
def test_power_spectrum_adjoint_matches_autograd(ctx, N):
    box = 80.0 / P.HUBBLE_H
    d, s, w = synth.snapshot_inputs(N, 6)
    dT = torch.from_numpy(3.0 + 2.0 * synth.from_padded(d)).cuda()
    bins, kmag, nb = ps_bins(N, box, 1.35)
    lam = np.random.default_rng(1).standard_normal(nb)
    g = reion.power_spectrum_adjoint(ctx, dT, box, lam).cpu().numpy()
    x = dT.double().clone().requires_grad_(True)
    p = ps_torch(x, box, 1.35, bins, kmag, nb)
    (p * torch.as_tensor(lam, device=x.device)).sum().backward()
    ref = x.grad.cpu().numpy()
    e = rel_l2(g, ref)
    print(f"N={N}: power-spectrum adjoint rel L2 vs float64 autograd {e:.2e}", flush=True)
    assert e < 1e-5


def _bubble_setup(N, L, z, eff, seed):
    d, s, w = synth.snapshot_inputs(N, seed)
    hp = P.hii_params(N, L, z, efficiency=eff, on_device=1)
    bp = P.tb_params(N, z, on_device=1)
    dd, ds, dw = (torch.from_numpy(a).cuda() for a in (d, s, w))
    return hp, bp, dd, ds, dw


def test_snapshot_adjoint_no_ts_matches_autograd(ctx):
    N, L, z = 64, 100.0, 8.0
    box = L / P.HUBBLE_H
    hp, bp, dd, ds, dw = _bubble_setup(N, L, z, 3.0, 11)
    dT, (xH, _, _, _) = MA.snapshot(ctx, None, hp, bp, dd, ds, dw)
    xh = xH.cpu().numpy().astype(np.float64)
    frac = ((xh > 0) & (xh < 1)).mean()
    print(f"partially ionised cells {frac:.3f}, ionised {(xh == 0).mean():.3f}", flush=True)
    assert frac > 0.05
    # reference power of a second realisation
    hp2, bp2, dd2, ds2, dw2 = _bubble_setup(N, L, z, 3.0, 12)
    dT2, _ = MA.snapshot(ctx, None, hp2, bp2, dd2, ds2, dw2)
    kf = MA.k_factor_for_bins(N, 12)
    Pref = MA.power_spectrum(ctx, dT2.detach(), box, kf).detach()
    # GPU: loss gradient through the hand-written adjoints
    dx = dd.clone().requires_grad_(True)
    sx = ds.clone().requires_grad_(True)
    dT, _ = MA.snapshot(ctx, None, hp, bp, dx, sx, dw)
    Pg = MA.power_spectrum(ctx, dT, box, kf)
    loss = ((Pg - Pref) ** 2).sum()
    loss.backward()
    gd = dx.grad.cpu().numpy()[:, :, :N]
    gs = sx.grad.cpu().numpy()[:, :, :N]
    # float64 autograd of the same map, mask (and xH outside the partial cells) from the GPU forward
    bins, kmag, nb = ps_bins(N, box, kf)
    dev = dd.device
    d64 = dd.double()[:, :, :N].clone().requires_grad_(True)
    s64 = ds.double()[:, :, :N].clone().requires_grad_(True)
    R = P.radius_ladder(hp)[-1]
    Vpix = (L / N) ** 3
    M = (4.0 / 3.0) * np.pi * R**3 * hp.omega_m * hp.rho_crit
    fc_unit = (4.0 / 3.0) * np.pi * R**3 / Vpix / M
    act = torch.as_tensor((xh > 0) & (xh < 1), device=dev)
    xh_t = torch.as_tensor(xh, device=dev)
    xpart = torch.clamp(1.0 - hp.efficiency * s64 * fc_unit / (1.0 + d64), 0.0, 1.0)
    xH64 = torch.where(act, xpart, xh_t)
    C = tb_const(z)
    dT64 = C * xH64 * (1.0 + d64)
    P64 = ps_torch(dT64, box, kf, bins, kmag, nb)
    L64 = ((P64 - Pref.double()) ** 2).sum()
    L64.backward()
    e_d = rel_l2(gd, d64.grad.cpu().numpy())
    e_s = rel_l2(gs, s64.grad.cpu().numpy())
    print(f"loss gpu {float(loss):.6e} float64 {float(L64):.6e}; grad rel L2: deltax {e_d:.2e} stars {e_s:.2e}",
          flush=True)
    assert abs(float(loss) - float(L64)) <= 1e-4 * abs(float(L64))
    assert e_d < 1e-3 and e_s < 1e-3


def _ts_case(N, L, seed, snap=24, nh=8):
    zz = P.snapshot_redshifts()
    z = float(zz[snap])
    d, s, w = synth.snapshot_inputs(N, seed)
    dd, ds, dw = (torch.from_numpy(a).cuda() for a in (d, s, w))
    hist = torch.stack([dw * float(((1 + z) / (1 + zz[max(snap - k, 0)])) ** 3) for k in range(nh)])
    fc = np.zeros(len(zz))
    fc[:snap] = np.linspace(1e-4, 5e-3, snap)
    tsp = P.ts_params(N, L, snap, zz, fc, nh, on_device=1)
    hp = P.hii_params(N, L, z, efficiency=8.0, include_spin_temp=1, on_device=1)
    bp = P.tb_params(N, z, include_spin_temp=1, on_device=1)
    xe0 = torch.full_like(dd, 2e-3)
    Tk0 = torch.full((N, N, N), 50.0, device=dd.device)
    return dict(zz=zz, fc=fc, tsp=tsp, hp=hp, bp=bp, d=dd, s=ds, w=dw, hist=hist, xe0=xe0, Tk0=Tk0)


def _ts_loss(ctx, c, d, s, Pref, box, kf, grad=False):
    if grad:
        d = d.clone().requires_grad_(True)
        s = s.clone().requires_grad_(True)
    dT, aux = MA.snapshot(ctx, c["tsp"], c["hp"], c["bp"], d, s, c["w"], c["w"], c["hist"], c["xe0"], c["Tk0"])
    Pg = MA.power_spectrum(ctx, dT, box, kf)
    loss = ((Pg - Pref) ** 2).sum()
    if grad:
        loss.backward()
        return loss, d.grad, s.grad, aux
    return float(loss), aux


def test_snapshot_adjoint_with_ts_finite_differences(ctx):
    # 32^3: a one-cell perturbation moves the float32 forward's loss by
    # ~eps / N^1.5 per mode against ~1e-7 of rounding, so FD is meaningful here
    N, L = 32, 25.0
    box = L / P.HUBBLE_H
    c = _ts_case(N, L, 21)
    # measure the tables' grid means once, then hold them fixed
    g0 = reion.compute_ts(ctx, c["tsp"], c["d"], c["s"], c["w"], c["hist"], c["xe0"].clone(), c["Tk0"].clone())
    c["tsp"].fixed_globals = 1
    c["tsp"].fixed_collapse_fraction = g0["collapse_fraction"]
    c["tsp"].fixed_x_e_ave = g0["x_e_ave"]
    assert g0["homogeneous"] == 0 and g0["no_light"] == 0
    c2 = _ts_case(N, L, 22)
    c2["tsp"] = c["tsp"]
    kf = MA.k_factor_for_bins(N, 12)
    dT2, _ = MA.snapshot(ctx, c2["tsp"], c2["hp"], c2["bp"], c2["d"], c2["s"], c2["w"], c2["w"], c2["hist"],
                         c2["xe0"], c2["Tk0"])
    Pref = MA.power_spectrum(ctx, dT2.detach(), box, kf).detach()
    loss, gd, gs, aux = _ts_loss(ctx, c, c["d"], c["s"], Pref, box, kf, grad=True)
    xH = aux[0].cpu().numpy()
    # threshold margins from the oracle's per-cell max of f_coll eff / n_f
    hpo = P.hii_params(N, L, c["hp"].redshift, efficiency=8.0, include_spin_temp=1)
    o = O.find_hii_bubbles(hpo, c["d"].cpu().numpy(), c["s"].cpu().numpy(), c["w"].cpu().numpy(),
                           x_e=aux[2].cpu().numpy(), want_ratio=True)
    margin = np.abs(o["max_ratio"].astype(np.float64) - 1.0)
    rng = np.random.default_rng(7)
    stars_n = c["s"].cpu().numpy()[:, :, :N]
    ok = (margin >= 1e-3) & (xH > 0) & (xH < 1)
    cand_d = np.argwhere(ok)
    # xH is affine in a cell's own stars, so the central difference is exact up
    # to rounding for any step that flips nothing: h = 1e-3 (internal mass
    # units, ~10% of a cell's stars) on cells 1e-2 away from the threshold
    cand_s = np.argwhere((margin >= 1e-2) & (xH > 0) & (xH < 1) & (stars_n > 2e-3))
    pick_d = cand_d[rng.choice(len(cand_d), size=64, replace=False)]
    pick_s = cand_s[rng.choice(len(cand_s), size=64, replace=False)]
    eps = 1e-3
    fd_d, ad_d, fd_s, ad_s = [], [], [], []
    gd_n = gd.cpu().numpy()
    gs_n = gs.cpu().numpy()
    for j in range(64):
        for which, store_fd, store_ad, gn, cell in (("d", fd_d, ad_d, gd_n, pick_d[j]),
                                                    ("s", fd_s, ad_s, gs_n, pick_s[j])):
            x, y, zc = (int(v) for v in cell)
            base = c[which]
            h = eps
            vals = []
            for sgn in (1.0, -1.0):
                pert = base.clone()
                pert[x, y, zc] += sgn * h
                dd_ = pert if which == "d" else c["d"]
                ss_ = pert if which == "s" else c["s"]
                lv, _ = _ts_loss(ctx, c, dd_, ss_, Pref, box, kf)
                vals.append(lv)
            store_fd.append((vals[0] - vals[1]) / (2 * h))
            store_ad.append(float(gn[x, y, zc]))
    fd_d, ad_d, fd_s, ad_s = map(np.array, (fd_d, ad_d, fd_s, ad_s))
    e_d = rel_l2(ad_d, fd_d)
    e_s = rel_l2(ad_s, fd_s)
    print(f"FD on 64 cells: deltax rel L2 {e_d:.2e} (median |rel| "
          f"{np.median(np.abs(ad_d - fd_d) / np.maximum(np.abs(fd_d), 1e-30)):.2e}), stars rel L2 {e_s:.2e}",
          flush=True)
    assert e_d < 5e-2 and e_s < 5e-2


def test_snapshot_adjoint_with_ts_matches_float64_autograd(ctx):
    """Every gradient of the spin-temperature snapshot against torch float64 autograd at 64^3.

    The reference (tests/ts_torch.py) rebuilds the shells from sfr and the SFR
    history, the per-cell Ts step, the partially ionised xH (the GPU forward's
    mask held fixed), delta_T and the power-spectrum loss in float64 from the
    oracle's snapshot constants, sharing no code with the HIP adjoint.
    """
    import ts_torch as TT

    N, L = 64, 50.0
    box = L / P.HUBBLE_H
    c = _ts_case(N, L, 31)
    g0 = reion.compute_ts(ctx, c["tsp"], c["d"], c["s"], c["w"], c["hist"], c["xe0"].clone(), c["Tk0"].clone())
    tsp = c["tsp"]
    tsp.fixed_globals = 1
    tsp.fixed_collapse_fraction = g0["collapse_fraction"]
    tsp.fixed_x_e_ave = g0["x_e_ave"]
    cst = O.ts_export(tsp)
    assert cst["nsteps"] == 40 and not cst["no_light"] and len(cst["ops"]) > 0
    c2 = _ts_case(N, L, 32)
    kf = MA.k_factor_for_bins(N, 12)
    dT2, _ = MA.snapshot(ctx, tsp, c2["hp"], c2["bp"], c2["d"], c2["s"], c2["w"], c2["w"], c2["hist"], c2["xe0"],
                         c2["Tk0"])
    Pref = MA.power_spectrum(ctx, dT2.detach(), box, kf).detach()
    # GPU: hand-written adjoint
    leaves = {k: v.clone().requires_grad_(True) for k, v in
              (("d", c["d"]), ("s", c["s"]), ("sfr", c["w"]), ("hist", c["hist"]), ("xe0", c["xe0"]),
               ("Tk0", c["Tk0"]))}
    dT, aux = MA.snapshot(ctx, tsp, c["hp"], c["bp"], leaves["d"], leaves["s"], c["w"], leaves["sfr"],
                          leaves["hist"], leaves["xe0"], leaves["Tk0"])
    loss = ((MA.power_spectrum(ctx, dT, box, kf) - Pref) ** 2).sum()
    loss.backward()
    xH_g, TS_g = aux[0].double(), aux[1].double()
    # float64 reference
    dev = c["d"].device
    r = {k: v.detach().double().clone().requires_grad_(True) for k, v in
         (("d", c["d"]), ("s", c["s"]), ("sfr", c["w"]), ("hist", c["hist"]), ("xe0", c["xe0"]), ("Tk0", c["Tk0"]))}
    h = tsp.hubble_h
    pixel_volume = (tsp.box_size / h / N) ** 3
    MT = tsp.unit_mass_g / tsp.unit_time_s
    L3 = tsp.unit_length_cm ** -3.0
    sh = TT.shells(cst, r["sfr"], r["hist"], N, tsp.box_size, dev)
    d64 = r["d"][:, :, :N]
    xn, Tn, TS = TT.ts_step(cst, d64, r["xe0"][:, :, :N], r["Tk0"], sh, pixel_volume, MT, L3)
    eTS = rel_l2(TS.detach().cpu().numpy(), TS_g.cpu().numpy())
    hp = c["hp"]
    R = P.radius_ladder(hp)[-1]
    fc_unit = (4.0 / 3.0) * np.pi * R**3 / (L / N) ** 3 / ((4.0 / 3.0) * np.pi * R**3 * hp.omega_m * hp.rho_crit)
    s64 = r["s"][:, :, :N]
    s64 = torch.where(s64 < 1e-8, torch.zeros_like(s64), s64)  # the stars clamp (ABS_TOL)
    xe_c = torch.where(xn < 1e-8, torch.zeros_like(xn), torch.clamp(xn, max=0.999))
    xpart = torch.clamp(1.0 - xe_c - hp.efficiency * s64 * fc_unit / (1.0 + d64), 0.0, 1.0)
    act = (xH_g > 0) & (xH_g < 1)
    xH64 = torch.where(act, xpart, xH_g)
    zf = float(np.float32(hp.redshift))
    dT64 = tb_const(hp.redshift) * xH64 * (1.0 + d64) * (1.0 - 2.728 * (1.0 + zf) / TS)
    bins, kmag, nb = ps_bins(N, box, kf)
    L64 = ((ps_torch(dT64, box, kf, bins, kmag, nb) - Pref.double()) ** 2).sum()
    L64.backward()
    print(f"partial cells {act.double().mean():.3f}; T_S float64 vs GPU rel L2 {eTS:.2e}; "
          f"loss gpu {float(loss.detach()):.6e} float64 {float(L64.detach()):.6e}", flush=True)
    assert eTS < 1e-5
    errs = {}
    for k in ("d", "s", "sfr", "hist", "xe0", "Tk0"):
        g = leaves[k].grad.detach().double().cpu().numpy()
        ref = r[k].grad.detach().cpu().numpy()
        errs[k] = rel_l2(g, ref)
    print("gradient rel L2 vs float64 autograd: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < 1e-4, k


def _fd_directions(ctx, c, Pref, box, kf, which, base, dirs, eps):
    """Central differences of the float64 power-spectrum loss along each direction."""
    out = []
    for v in dirs:
        vals = []
        for sgn in (1.0, -1.0):
            pert = base + sgn * eps * v
            dd_ = pert if which == "d" else c["d"]
            ss_ = pert if which == "s" else c["s"]
            vals.append(_ts_loss(ctx, c, dd_, ss_, Pref, box, kf)[0])
        out.append((vals[0] - vals[1]) / (2 * eps))
    return np.array(out)


def test_snapshot_adjoint_fd_directions_256(ctx):
    """The adjoint configuration at 256^3 (200 Mpc/h, the 512^3 cell size): the
    hand-written gradient of the power-spectrum loss against central finite
    differences of the GPU forward along 64 random directions in deltax and 64
    in stars.  A deltax direction is +-1 on 8192 partially ionised cells at
    least 1e-3 from the threshold, step 1e-3.  A stars direction is +-(the
    cell's own stars) on 2048 partially ionised cells at least 0.1 from it, step
    0.05: a partially ionised cell's xH is affine in its own stars and the loss
    quadratic in xH, so the central difference is exact up to rounding for any
    step that flips nothing, and a step of 5 % moves a cell's ratio by at most
    5 % of itself, below its margin.  (A step of 1e-3 of the stars on 512 cells
    moves the loss by less than the float32 forward's rounding at this size.)
    The loss is summed in float64; the shell tables' grid means are held fixed
    in both."""
    N, L = 256, 200.0
    box = L / P.HUBBLE_H
    print("adjoint FD at 256^3: forward, backward and oracle margins", flush=True)
    c = _ts_case(N, L, 31)
    g0 = reion.compute_ts(ctx, c["tsp"], c["d"], c["s"], c["w"], c["hist"], c["xe0"].clone(), c["Tk0"].clone())
    c["tsp"].fixed_globals = 1
    c["tsp"].fixed_collapse_fraction = g0["collapse_fraction"]
    c["tsp"].fixed_x_e_ave = g0["x_e_ave"]
    c2 = _ts_case(N, L, 32)
    kf = MA.k_factor_for_bins(N, 20)
    dT2, _ = MA.snapshot(ctx, c["tsp"], c2["hp"], c2["bp"], c2["d"], c2["s"], c2["w"], c2["w"], c2["hist"],
                         c2["xe0"], c2["Tk0"])
    Pref = MA.power_spectrum(ctx, dT2.detach(), box, kf).detach()
    del c2, dT2
    loss, gd, gs, aux = _ts_loss(ctx, c, c["d"], c["s"], Pref, box, kf, grad=True)
    xH = aux[0].cpu().numpy()
    hpo = P.hii_params(N, L, c["hp"].redshift, efficiency=8.0, include_spin_temp=1)
    o = O.find_hii_bubbles(hpo, c["d"].cpu().numpy(), c["s"].cpu().numpy(), c["w"].cpu().numpy(),
                           x_e=aux[2].cpu().numpy(), want_ratio=True)
    margin = np.abs(o["max_ratio"].astype(np.float64) - 1.0)
    partial = (xH > 0) & (xH < 1)
    stars_n = c["s"].cpu().numpy()[:, :, :N]
    cand_d = np.argwhere(partial & (margin >= 1e-3))
    cand_s = np.argwhere(partial & (margin >= 0.1) & (stars_n > 0))
    rng = np.random.default_rng(17)
    res = {}
    for which, cand, K, eps, g in (("d", cand_d, 8192, 1e-3, gd), ("s", cand_s, 2048, 0.05, gs)):
        dirs = []
        for _ in range(64):
            cells = cand[rng.choice(len(cand), size=K, replace=False)]
            v = torch.zeros_like(c[which])
            sg = torch.from_numpy(rng.choice([-1.0, 1.0], size=K).astype(np.float32)).to(v.device)
            idx = (cells[:, 0], cells[:, 1], cells[:, 2])
            v[idx] = sg if which == "d" else sg * c[which][idx]
            dirs.append(v)
        fd = _fd_directions(ctx, c, Pref, box, kf, which, c[which], dirs, eps)
        ad = np.array([float((g * v).sum()) for v in dirs])
        res[which] = rel_l2(ad, fd)
        print(f"  {which}: 64 directions of {K} cells, step {eps:g}, AD vs FD rel L2 {res[which]:.2e}, median |rel| "
              f"{np.median(np.abs(ad - fd) / np.abs(fd)):.2e}", flush=True)
    assert res["d"] < 2e-2 and res["s"] < 2e-2


def test_fixed_table_means_term_64(ctx):
    """The shell tables' two grid means (collapse fraction and <x_e>,
    ComputeTs.c:413-475, which the tables depend on through tau_X), along 32
    deltax directions at 64^3: central differences of the loss with the means
    held fixed and with the forward measuring them again differ by about 3 %
    (bounded here at 5 %); the adjoint with the means fixed matches the first,
    and the adjoint with them measured (its gradient through the tables'
    derivatives with respect to the means) matches the second."""
    N, L = 64, 50.0
    box = L / P.HUBBLE_H
    c = _ts_case(N, L, 31)
    g0 = reion.compute_ts(ctx, c["tsp"], c["d"], c["s"], c["w"], c["hist"], c["xe0"].clone(), c["Tk0"].clone())
    free = c["tsp"]
    fixed = P.ts_params(N, L, 24, c["zz"], c["fc"].copy(), 8, on_device=1)
    fixed.fixed_globals = 1
    fixed.fixed_collapse_fraction = g0["collapse_fraction"]
    fixed.fixed_x_e_ave = g0["x_e_ave"]
    c2 = _ts_case(N, L, 32)
    kf = MA.k_factor_for_bins(N, 12)
    dT2, _ = MA.snapshot(ctx, fixed, c2["hp"], c2["bp"], c2["d"], c2["s"], c2["w"], c2["w"], c2["hist"],
                         c2["xe0"], c2["Tk0"])
    Pref = MA.power_spectrum(ctx, dT2.detach(), box, kf).detach()
    c["tsp"] = fixed
    loss, gd, gs, aux = _ts_loss(ctx, c, c["d"], c["s"], Pref, box, kf, grad=True)
    xH = aux[0].cpu().numpy()
    hpo = P.hii_params(N, L, c["hp"].redshift, efficiency=8.0, include_spin_temp=1)
    o = O.find_hii_bubbles(hpo, c["d"].cpu().numpy(), c["s"].cpu().numpy(), c["w"].cpu().numpy(),
                           x_e=aux[2].cpu().numpy(), want_ratio=True)
    margin = np.abs(o["max_ratio"].astype(np.float64) - 1.0)
    cand = np.argwhere((xH > 0) & (xH < 1) & (margin >= 1e-3))
    rng = np.random.default_rng(3)
    dirs = []
    for _ in range(32):
        cells = cand[rng.choice(len(cand), size=64, replace=False)]
        v = torch.zeros_like(c["d"])
        v[cells[:, 0], cells[:, 1], cells[:, 2]] = torch.from_numpy(
            rng.choice([-1.0, 1.0], size=64).astype(np.float32)).to(v.device)
        dirs.append(v)
    fd_fixed = _fd_directions(ctx, c, Pref, box, kf, "d", c["d"], dirs, 1e-3)
    c["tsp"] = free
    fd_free = _fd_directions(ctx, c, Pref, box, kf, "d", c["d"], dirs, 1e-3)
    ad = np.array([float((gd * v).sum()) for v in dirs])
    dropped = rel_l2(fd_fixed, fd_free)
    _, gd_free, _, _ = _ts_loss(ctx, c, c["d"], c["s"], Pref, box, kf, grad=True)
    ad_free = np.array([float((gd_free * v).sum()) for v in dirs])
    print(f"fixed table means: FD fixed vs measured rel L2 {dropped:.2e}; AD (fixed) vs FD fixed "
          f"{rel_l2(ad, fd_fixed):.2e}, vs FD measured {rel_l2(ad, fd_free):.2e}; AD (measured) vs FD measured "
          f"{rel_l2(ad_free, fd_free):.2e}", flush=True)
    assert rel_l2(ad, fd_fixed) < 2e-2
    assert dropped < 5e-2
    assert rel_l2(ad_free, fd_free) < 2e-2


def test_table_means_gradient_xe_64(ctx):
    """<x_e>, one of the two grid means the shell tables are built from, is the
    mean of x_e_prev: along the uniform direction in x_e_prev (every cell moved
    by the same step, so <x_e> moves by it too) the adjoint with the means
    measured matches central differences of the forward that measures them
    again, and the adjoint with them held fixed matches the forward that holds
    them fixed; the two derivatives differ by the tables' share, about 1 %
    here, which the adjoint reproduces (64.5 against 61.7 measured)."""
    N, L = 64, 50.0
    box = L / P.HUBBLE_H
    c = _ts_case(N, L, 41)
    g0 = reion.compute_ts(ctx, c["tsp"], c["d"], c["s"], c["w"], c["hist"], c["xe0"].clone(), c["Tk0"].clone())
    free = c["tsp"]
    fixed = P.ts_params(N, L, 24, c["zz"], c["fc"].copy(), 8, on_device=1)
    fixed.fixed_globals = 1
    fixed.fixed_collapse_fraction = g0["collapse_fraction"]
    fixed.fixed_x_e_ave = g0["x_e_ave"]
    c2 = _ts_case(N, L, 42)
    kf = MA.k_factor_for_bins(N, 12)
    dT2, _ = MA.snapshot(ctx, fixed, c2["hp"], c2["bp"], c2["d"], c2["s"], c2["w"], c2["w"], c2["hist"],
                         c2["xe0"], c2["Tk0"])
    Pref = MA.power_spectrum(ctx, dT2.detach(), box, kf).detach()
    real = torch.zeros_like(c["xe0"])
    real[:, :, :N] = 1.0

    def loss(tsp, xe, grad=False):
        if grad:
            xe = xe.clone().requires_grad_(True)
        dT, _ = MA.snapshot(ctx, tsp, c["hp"], c["bp"], c["d"], c["s"], c["w"], c["w"], c["hist"], xe, c["Tk0"])
        L_ = ((MA.power_spectrum(ctx, dT, box, kf) - Pref) ** 2).sum()
        if grad:
            L_.backward()
            return float((xe.grad * real).sum())
        return float(L_)

    eps = 2e-6  # 1e-3 of x_e_prev = 2e-3
    out = {}
    for name, tsp in (("fixed", fixed), ("measured", free)):
        fd = (loss(tsp, c["xe0"] + eps * real) - loss(tsp, c["xe0"] - eps * real)) / (2 * eps)
        ad = loss(tsp, c["xe0"], grad=True)
        out[name] = (ad, fd)
    print(f"uniform x_e_prev direction: means fixed AD {out['fixed'][0]:.6e} FD {out['fixed'][1]:.6e}; "
          f"measured AD {out['measured'][0]:.6e} FD {out['measured'][1]:.6e}", flush=True)
    for ad, fd in out.values():
        assert abs(ad - fd) <= 2e-2 * abs(fd)
    # the tables' share (about 1 % of the derivative at this snapshot): the
    # adjoint's matches the finite differences' to 20 % of itself
    share_fd = out["measured"][1] - out["fixed"][1]
    share_ad = out["measured"][0] - out["fixed"][0]
    print(f"  the tables' share: AD {share_ad:.3e} FD {share_fd:.3e}", flush=True)
    assert abs(share_fd) > 2e-3 * abs(out["fixed"][1])
    assert abs(share_ad - share_fd) <= 0.2 * abs(share_fd)
