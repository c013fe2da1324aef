"""Benchmark of the excursion-set reionization path on one MI355X.

A step is one snapshot of the hot path on a 512^3, 400 Mpc/h box, all inputs
resident in HBM when the timed region starts:
  --workload snapshot (default): the reference's order per snapshot - the
      spin-temperature step (40 shells, inhomogeneous branch, SFR history),
      find_HII_bubbles with x_e as the fourth field (44 radii, top-hat, UVB on),
      then the brightness-temperature grid with T_S;
  --workload bubbles: find_HII_bubbles (three fields) + brightness temperature.
The metric is cells*radii per second (N^3 * bubble radii / step time).

Besides the headline line it reports, from HIP events recorded around every
kernel launch on the stream the kernels run on, the dominant kernel's average
launch time and its achieved HBM bandwidth against the MI355X peak (and its
PMC-measured bytes per launch where a summary under profiles/ holds them), the
per-kernel table, the device time per stage, the path's fraction of peak on
the 3-pass byte model, a 2 GiB device-copy ceiling, and a CPU baseline: the
oracle (the C restatement of the reference's FFTW-float path) on the host cores
over a bounded sample of the same workload.  Unless --no-extras is given the
snapshot run also times the hand-written adjoint against its forward at 256^3
and 512^3, the batched bubble finder (96 uneven boxes, one graph launch)
against the per-box loop, the galaxy gridding (2^24 galaxies onto 512^3),
delta_T with peculiar velocities and line-of-sight RSD at 512^3, and the HDF5
grid files of one snapshot written from HBM at 256^3.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 512]
           [--workload snapshot|bubbles|batched|gridding|grid_files] [--no-extras] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def model_bytes_hii(N, F, S, uvb):
    """3-pass algorithmic byte model of the bubble finder (bytes per snapshot)."""
    Bc = 8 * N * N * (N // 2 + 1)
    Nr = N**3
    W = 20 + (8 if uvb else 0)
    return F * 6 * Bc + S * (F * 6 * Bc + (4 * F + W) * Nr)


def kernel_bytes(name, N, F, radii=1.0):
    """Algorithmic HBM bytes of one launch of a kernel (packed spectra: 8*N*N*N/2 bytes).

    Names with a ts_ prefix are launches of the spin-temperature shell loop,
    one field each; the others are the bubble finder's, F fields each.
    ``radii`` is the average number of ladder radii one ix_filter_k launch
    covers (it reads each field's spectrum once and writes one per radius).
    """
    if name.startswith("ts_"):
        name, F = name[3:], 1
        if name in ("ix_filter", "colpass_y_inv", "c2r_z", "colpass_y_fwd", "colpass_x_fwd"):
            # sharp-k shells: only the tiles holding the ball of kept modes are
            # read (and, in the x/y passes, written), and the forward passes of
            # the working grid transform only the tiles the shells reusing it
            # read, so the bytes depend on R
            return None
        if name == "cell":
            name = "ts_cell"
    spec = 8 * N * N * (N // 2)
    Nr = N**3
    table = {
        # strided pencil passes: read + write one spectrum per field
        "colpass_y_inv": 2 * F * spec,
        "colpass_y_fwd": 2 * spec,  # forward passes run one field per launch
        "ts_cell": 40 * 4 * Nr + 4 * N * N * (N + 2) * 3 + 3 * 4 * Nr,
        "colpass_x_fwd": 2 * spec,
        # window + inverse x pass: read unfiltered (+ Nyquist plane) and write per field
        "ix_filter": F * (2 * spec + 8 * N * N),
        "ix_filter_k": F * ((1 + radii) * spec + 8 * N * N),
        # inverse z pass + threshold: read each field's spectrum, 1 byte of step state per cell
        "c2r_threshold": F * spec + 2 * Nr,
        "r2c_z": 4 * N * N * (N + 2) + spec,
        "window_table": 4 * (N // 2 + 1) ** 3,
        "brightness_temp": 4 * N * N * (N + 2) + 2 * 4 * Nr,
    }
    return table.get(name)


def path_bytes_hii(N, F, S):
    """The bubble finder's algorithmic bytes for one box, the sum of its kernels'
    (kernel_bytes): F forward transforms, then per radius the windowed inverse x
    pass, the inverse y pass and the z pass + threshold; the window tables
    (shared by the boxes of a batch cohort) are left out."""
    spec = 8 * N * N * (N // 2)
    fwd = F * (kernel_bytes("r2c_z", N, F) + 2 * 2 * spec)
    per_radius = kernel_bytes("ix_filter", N, F) + kernel_bytes("colpass_y_inv", N, F) + kernel_bytes(
        "c2r_threshold", N, F)
    return fwd + S * per_radius


def cpu_info():
    """The host's CPU model, its logical CPUs and the CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "logical_cpus": os.cpu_count()}


def _cpu_snapshot(n, box, threads):
    """One snapshot of the oracle (Ts + bubbles with x_e + dTb) on an n^3 box; seconds by part."""
    import numpy as np

    from meraxes_amd import params as P
    from meraxes_amd.snapshots import SyntheticRun
    from oracle import oracle as O

    O.set_threads(threads)
    zz = P.snapshot_redshifts()
    snap = 24
    z = float(zz[snap])
    d, s, w = SyntheticRun(n, box, seed=1).inputs(snap)
    nh = 8
    hist = np.stack([w * np.float32(((1 + z) / (1 + zz[max(snap - k, 0)])) ** 3) for k in range(nh)])
    fcoll = np.zeros(len(zz))
    fcoll[:snap] = np.linspace(1e-4, 5e-3, snap)
    tsp = P.ts_params(n, box, snap, zz, fcoll, nh, efficiency=20.0)
    xe0 = np.full_like(d, 2e-3)
    Tk0 = np.full((n, n, n), 50.0, np.float32)
    p = P.hii_params(n, box, z, filter_type=0, efficiency=20.0, uvb_flag=1, include_spin_temp=1)
    tp = P.tb_params(n, z, include_spin_temp=1)
    t0 = time.perf_counter()
    ts = O.compute_ts(tsp, d, s, w, hist, xe0, Tk0)
    t1 = time.perf_counter()
    r = O.find_hii_bubbles(p, d, s, w, x_e=ts["x_e_box"])
    O.brightness_temp(tp, r["xH"], d, ts["TS_box"])
    t2 = time.perf_counter()
    return r["n_radii"], t2 - t0, t1 - t0, O.get_threads()


def cpu_baseline(snapshot):
    """The oracle (ref32 restatement of the reference's CPU path) on bounded samples.

    snapshot: spin temperature + bubble finder with x_e + brightness
    temperature of z = 8.41 on boxes with the 512^3 / 400 Mpc/h cell size:
    160^3 (125 Mpc/h) on every host core the process may use, and 96^3
    (75 Mpc/h) on one thread (the reference runs one thread per MPI rank).
    bubbles: find_HII_bubbles over the full ladder of a 256^3, 200 Mpc/h box.
    """
    from meraxes_amd import params as P
    from meraxes_amd import synth
    from oracle import oracle as O

    info = cpu_info()
    if not snapshot:
        n, box = 256, 200.0
        d, s, w = synth.snapshot_inputs(n, 1)
        p = P.hii_params(n, box, 8.0, filter_type=0, efficiency=20.0, uvb_flag=1)
        O.set_threads(info["nproc"])
        t0 = time.perf_counter()
        r = O.find_hii_bubbles(p, d, s, w)
        dt = time.perf_counter() - t0
        S = r["n_radii"]
        return dict(value=n**3 * S / dt, unit="cells*radii/s", cores=O.get_threads(), kind="port",
                    sample=f"oracle ref32: find_HII_bubbles, {n}^3 box {box:g} Mpc/h z=8, S={S}, F=3; {dt:.1f} s",
                    **info)
    n, box = 160, 125.0
    S, dt, dts, threads = _cpu_snapshot(n, box, info["nproc"])
    n1, box1 = 96, 75.0
    S1, dt1, dts1, _ = _cpu_snapshot(n1, box1, 1)
    return {
        "value": n**3 * S / dt,
        "unit": "cells*radii/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle ref32, ComputeTs (40 shells) + find_HII_bubbles (F=4, S={S}) + dTb at z=8.414, "
                   f"{n}^3 box {box:g} Mpc/h, {threads} threads: {dt:.1f} s (Ts {dts:.1f} s, of which the "
                   f"single-threaded numpy host tables about 2.3 s)"),
        "single_thread": {"value": n1**3 * S1 / dt1, "unit": "cells*radii/s", "cores": 1,
                          "sample": f"the same step on a {n1}^3, {box1:g} Mpc/h box (S={S1}), one thread: "
                                    f"{dt1:.1f} s (Ts {dts1:.1f} s)"},
        **info,
    }


def copy_ceiling(torch, ctx, nbytes=2 << 30, reps=10):
    """Practical HBM ceilings for a 2 GiB device-to-device copy (bytes read + written
    per second): the library's streaming copy kernel (one 16-byte load and one
    non-temporal store per thread over a grid covering the buffer; HIP events on
    its stream) and torch's copy_."""
    import ctypes as C

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    ms = C.c_double()
    ctx.check(ctx.lib.mrx_stream_copy(ctx.ptr, C.c_void_p(dst.data_ptr()), C.c_void_p(src.data_ptr()), nbytes, reps,
                                      C.byref(ms)), "stream_copy")
    assert torch.equal(dst[:: 1 << 16], src[:: 1 << 16])
    t_stream = ms.value / 1e3
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    del src, dst
    return {"GB_s": 2 * nbytes / t_stream / 1e9, "frac_of_peak": 2 * nbytes / t_stream / HBM_PEAK, "bytes": nbytes,
            "kernel": "mrx_stream_copy (2 GiB, 10 repetitions)",
            "torch_copy_GB_s": 2 * nbytes / t / 1e9, "torch_copy_frac_of_peak": 2 * nbytes / t / HBM_PEAK}


def measured_traffic(kernel, N):
    """HBM bytes per launch of `kernel` from the committed PMC summary, or None.

    profiles/traffic_<kernel>_<N>.json holds FETCH_SIZE + WRITE_SIZE per
    dispatch from rocprofv3 --pmc, scaled by the same counters' reading of a
    copy kernel of known size (see the summary beside it).
    """
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{N}.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def bench_adjoint(ctx, N, steps=2):
    """Forward (Ts + bubbles + dTb + power-spectrum loss) vs its hand-written backward at N^3.

    The loss is the squared difference of the binned 21-cm power (20 bins) to
    that of a second seeded realisation; gradients reach deltax and stars.
    """
    import numpy as np
    import torch

    from meraxes_amd import autograd as MA
    from meraxes_amd import params as P
    from meraxes_amd import synth

    dev = torch.device("cuda:0")
    L = 400.0 * N / 512
    zz = P.snapshot_redshifts()
    snap = 24
    z = float(zz[snap])
    nh = 8
    log(f"[bench] adjoint at {N}^3: generating inputs")

    def case(seed):
        d, s_, w = synth.snapshot_inputs_torch(N, seed, dev)
        hist = torch.stack([w * float(((1 + z) / (1 + zz[max(snap - k, 0)])) ** 3) for k in range(nh)])
        return d, s_, w, hist

    fc = np.zeros(len(zz))
    fc[:snap] = np.linspace(1e-4, 5e-3, snap)
    tsp = P.ts_params(N, L, snap, zz, fc, nh, efficiency=20.0, on_device=1)
    hp = P.hii_params(N, L, z, filter_type=0, efficiency=20.0, uvb_flag=0, include_spin_temp=1, on_device=1)
    bp = P.tb_params(N, z, include_spin_temp=1, on_device=1)
    xe0 = None
    Tk0 = None
    box = L / P.HUBBLE_H
    kf = MA.k_factor_for_bins(N, 20)
    d2, s2, w2, h2 = case(1)
    xe0 = torch.full_like(d2, 2e-3)
    Tk0 = torch.full((N, N, N), 50.0, device=dev)
    dT2, _ = MA.snapshot(ctx, tsp, hp, bp, d2, s2, w2, w2, h2, xe0, Tk0)
    Pref = MA.power_spectrum(ctx, dT2.detach(), box, kf).detach()
    del d2, s2, w2, h2, dT2
    d, s_, w, hist = case(0)
    t_f, t_b = [], []
    for it in range(steps + 1):
        dx = d.clone().requires_grad_(True)
        sx = s_.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dT, _ = MA.snapshot(ctx, tsp, hp, bp, dx, sx, w, w, hist, xe0, Tk0)
        Pk = MA.power_spectrum(ctx, dT, box, kf)
        loss = ((Pk - Pref) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it > 0:  # the first pass warms up
            t_f.append(t1 - t0)
            t_b.append(t2 - t1)
        g_norm = float(dx.grad.double().norm())
        del dx, sx, dT, Pk, loss
    tf, tb = float(np.median(t_f)), float(np.median(t_b))
    log(f"[bench] adjoint {N}^3: forward {tf * 1e3:.1f} ms, backward {tb * 1e3:.1f} ms, ratio {tb / tf:.2f} "
        f"(|dL/d deltax| = {g_norm:.3e})")
    return {"forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "backward_over_forward": tb / tf,
            "loss": "sum over 20 log-k bins of (P_dTb - P_ref)^2, P_ref from a second seed",
            "grads": ["deltax", "stars"]}


BATCH_SIZES = (64, 128, 192)
def bench_gridding(ctx, N=512, n_gal=1 << 24, reps=3):
    """Galaxy gridding (construct_baryon_grids) of 2^24 galaxies onto 512^3 grids.

    Spin temperature on (stars, weighted_sfr, sfr and an 8-slot SFR history
    ring), black-hole feedback on; galaxy arrays and grids resident in HBM.
    Positions uniform with a tenth of the galaxies in 16 clumps.
    """
    import numpy as np
    import torch

    from meraxes_amd import params as P
    from meraxes_amd import reion

    L = 400.0 * N / 512
    g = torch.Generator(device="cuda").manual_seed(11)
    dev = torch.device("cuda:0")
    pos = torch.rand((n_gal, 3), generator=g, device=dev) * L
    k = n_gal // 10
    centres = torch.rand((16, 3), generator=g, device=dev) * L
    pick = torch.randint(0, 16, (k,), generator=g, device=dev)
    pos[:k] = torch.remainder(centres[pick] + torch.randn((k, 3), generator=g, device=dev) * (L / 400), L)
    pos = pos.float().clamp_(0, float(np.nextafter(np.float32(L), 0))).contiguous()
    type_ = torch.randint(0, 4, (n_gal,), generator=g, device=dev, dtype=torch.int32)
    logn = (lambda m, s: torch.exp(m + s * torch.randn(n_gal, generator=g, device=dev, dtype=torch.float64)))
    fesc, gross, bhm, ebhm = logn(-3, 2), logn(-2, 2), logn(-4, 1.5), logn(-6, 2)
    p = P.grid_params(N, L, 8.4, include_spin_temp=1, n_store_sfr=8, bh_feedback=1)
    shape = (N, N, 2 * (N // 2 + 1))
    stars, wsfr, sfr = (torch.empty(shape, device=dev) for _ in range(3))
    hist = torch.zeros((8,) + shape, device=dev)
    run = (lambda: reion.construct_baryon_grids(ctx, p, pos, type_, fesc, gross, bhm, ebhm, stars=stars,
                                                 weighted_sfr=wsfr, sfr=sfr, sfr_histories=hist))
    r = run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    np_ = 2 * N * N * (N // 2 + 1)
    # per call: galaxy arrays (12 + 4 + 4 * 8 B) read for the keys and the sums,
    # keys/values written, sorted and read; grids zeroed and written; the
    # history ring moved by one slot (7 of 8 slots read and written)
    grid_bytes = 4 * np_ * (3 + 3 + 2 * 7 + 1)
    log(f"[bench] gridding {n_gal} galaxies onto {N}^3: {dt * 1e3:.2f} ms ({n_gal / dt:.3e} galaxies/s)")
    del pos, type_, fesc, gross, bhm, ebhm, stars, wsfr, sfr, hist
    torch.cuda.empty_cache()
    return {"metric": "galaxies/s of construct_baryon_grids (stars, weighted_sfr, sfr, 8-slot history)",
            "value": n_gal / dt, "ms_per_call": dt * 1e3, "n_galaxies": n_gal, "grid": N,
            "n_gridded": r["n_gridded"], "grid_bytes_per_call": grid_bytes,
            "grid_GB_s": grid_bytes / dt / 1e9,
            "config": f"{n_gal} galaxies, {N}^3, {L:g} Mpc/h, spin temperature and black-hole feedback on, "
                      f"a tenth in 16 clumps, synthetic (torch generator)"}


def bench_tb_vel(ctx, N=512, reps=3):
    """delta_T with peculiar velocities (Flag_IncludePecVelsFor21cm = 3) at 512^3.

    Velocity conversion, FFT gradient, optical-depth delta_T and the
    line-of-sight sub-cell re-binning (50 sub-cells per cell); a smooth random
    velocity field of 0.25 cell rms displacement, grids resident in HBM.
    """
    import math

    import torch

    from meraxes_amd import params as P
    from meraxes_amd import reion

    L = 400.0 * N / 512
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cuda").manual_seed(21)
    xH = torch.rand((N, N, N), generator=g, device=dev)
    deltax = torch.randn((N, N, N + 2), generator=g, device=dev) * 0.3
    TS = torch.rand((N, N, N), generator=g, device=dev) * 200 + 15
    z = 8.0
    Hz = 3.2407789e-18 * P.HUBBLE_H * math.sqrt(P.OMEGA_M * (1 + z) ** 3 + P.OMEGA_L)
    unit = (1 + z) * 1e5 / 3.086e24 * (1 + z)
    # smooth field: box-filtered noise along z, displacements of 0.25 cell rms
    # (a 30 km/s peculiar velocity at z = 8 moves gas by about 0.1-0.3 cells here)
    v0 = torch.randn((N, N, N + 2), generator=g, device=dev)
    v0 = (v0 + torch.roll(v0, 1, 2) + torch.roll(v0, -1, 2) + torch.roll(v0, 2, 2)) / 8.0
    v0 *= float((L / P.HUBBLE_H) / N * Hz / unit)
    vel = torch.empty_like(v0)
    p = P.tb_vel_params(N, L, z, pec_vels=3, trees_id=0)
    out = torch.empty((N, N, N), device=dev)
    vel.copy_(v0)
    reion.brightness_temp_vel(ctx, p, xH, deltax, TS, vel, out=out)
    torch.cuda.synchronize()
    t = 0.0
    for _ in range(reps):
        vel.copy_(v0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reion.brightness_temp_vel(ctx, p, xH, deltax, TS, vel, out=out)
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
    dt = t / reps
    log(f"[bench] delta_T with velocities (RSD) at {N}^3: {dt * 1e3:.2f} ms")
    del xH, deltax, TS, v0, vel, out
    torch.cuda.empty_cache()
    return {"ms_per_call": dt * 1e3, "cells_per_s": N**3 / dt, "grid": N, "pec_vels": 3,
            "config": f"{N}^3, {L:g} Mpc/h, z=8, spin temperature, RSD with 50 sub-cells, synthetic"}


def bench_grid_files(ctx, N=256, reps=2):
    """The HDF5 grid files (save_reion_input_grids + save_reion_output_grids) from HBM.

    Every flag on: 4 input grids (3 padded and converted, sfr and weighted_sfr
    to M_sun/yr) and 10 output grids (x_e_box de-padded), N^3 float32 each,
    written to a temporary directory on the local disk and deleted after each
    repetition.  Reported against two ceilings measured in the same run: a
    device-to-pinned-host copy of one grid, and a plain write() of the same
    bytes to a file in the same directory (no fsync, as HDF5 does not sync).
    """
    import shutil
    import tempfile

    import numpy as np
    import torch

    from meraxes_amd import abi, reion
    from meraxes_amd import params as P

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cuda").manual_seed(5)
    pad = lambda: torch.rand((N, N, N + 2), generator=g, device=dev)  # noqa: E731
    real = lambda: torch.rand((N, N, N), generator=g, device=dev)  # noqa: E731
    inputs = [pad() for _ in range(4)]
    grids = {n: real() for n in ("xH", "z_at_ionization", "r_bubble", "J_21", "J_21_at_ionization", "Mvir_crit",
                                 "TS_box", "Tk_box", "delta_T")}
    grids["x_e_box_prev"] = pad()
    avg = {n: 0.5 for n in abi.OUTPUT_AVERAGES}
    p = P.save_params(N, uvb_flag=1, include_spin_temp=1, compute_21cm_bright_temp=1)
    d = tempfile.mkdtemp(prefix="mrx_grids_", dir=os.environ.get("TMPDIR", "/tmp"))
    path = os.path.join(d, "meraxes_grids_24.hdf5")
    nbytes = 14 * 4 * N**3
    try:
        def once():
            reion.save_reion_input_grids(ctx, p, path, *inputs)
            reion.save_reion_output_grids(ctx, p, path, grids, avg)

        torch.cuda.synchronize()
        once()
        os.remove(path)
        t = 0.0
        for _ in range(reps):
            t0 = time.perf_counter()
            once()
            t += time.perf_counter() - t0
            os.remove(path)
        dt = t / reps
        # ceilings: one grid device -> pinned host, and the same bytes through write()
        host = torch.empty((N, N, N), dtype=torch.float32, pin_memory=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(4):
            host.copy_(grids["xH"])
        torch.cuda.synchronize()
        d2h = 4 * 4 * N**3 / (time.perf_counter() - t0)
        arr = host.numpy()
        raw = os.path.join(d, "raw.bin")
        t0 = time.perf_counter()
        with open(raw, "wb") as f:
            for _ in range(14):
                arr.tofile(f)
        wr = nbytes / (time.perf_counter() - t0)
        os.remove(raw)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    log(f"[bench] grid files at {N}^3: {dt * 1e3:.1f} ms for {nbytes / 2**20:.0f} MiB "
        f"({nbytes / dt / 1e9:.2f} GB/s; D2H ceiling {d2h / 1e9:.1f} GB/s, write() ceiling {wr / 1e9:.2f} GB/s)")
    del inputs, grids
    torch.cuda.empty_cache()
    return {"ms_per_snapshot": dt * 1e3, "GB_s": nbytes / dt / 1e9, "bytes": nbytes, "grid": N,
            "d2h_ceiling_GB_s": d2h / 1e9, "write_ceiling_GB_s": wr / 1e9,
            "frac_of_write_ceiling": (nbytes / dt) / wr,
            "config": f"{N}^3, all flags: 4 input + 10 output grids from HBM to one HDF5 file, synthetic"}


BATCH_PER_SIZE = 32


def bench_batched(ctx, steps, warmup, per_size=BATCH_PER_SIZE):
    """96 uneven boxes (32 each of 64^3, 128^3, 192^3, 100 Mpc/h, z=8, top-hat,
    three fields, efficiency U[10, 40], own seed each): the batch replayed as one
    graph launch, against the per-box plain loop over the same boxes."""
    import numpy as np
    import torch

    from meraxes_amd import params as P
    from meraxes_amd import reion, synth

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(4)
    boxes = []
    cells_radii = 0
    mbytes = 0
    log(f"[bench] batched: generating {per_size * len(BATCH_SIZES)} boxes on the GPU")
    for N in BATCH_SIZES:
        for i in range(per_size):
            seed = 1000 * N + i
            d, s, w = synth.snapshot_inputs_torch(N, seed, dev)
            p = P.hii_params(N, 100.0, 8.0, filter_type=0, efficiency=float(rng.uniform(10.0, 40.0)))
            S = len(P.radius_ladder(p))
            cells_radii += N**3 * S
            mbytes += model_bytes_hii(N, 3, S, False)
            boxes.append({"params": p, "deltax": d, "stars": s, "weighted_sfr": w})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch = reion.HiiBatch(ctx, boxes, graph=True, tune=True)
    t_plan = time.perf_counter() - t0
    info = batch.describe()
    log(f"[bench] batched plan {t_plan:.2f} s: {info}")
    for _ in range(warmup):
        batch.run(sync=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run(sync=False)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    log(f"[bench] batched: {t_batch * 1e3:.2f} ms per graph launch")
    # the per-box plain loop over the same boxes and outputs
    for N in BATCH_SIZES:
        b = next(x for x in batch.boxes if x["params"].grid_dim == N)
        reion.find_hii_bubbles(ctx, b["params"], b["deltax"], b["stars"], b["weighted_sfr"],
                               z_at_ionization=b["z_at_ionization"], out=b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in batch.boxes:
        reion.find_hii_bubbles(ctx, b["params"], b["deltax"], b["stars"], b["weighted_sfr"],
                               z_at_ionization=b["z_at_ionization"], out=b)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    log(f"[bench] per-box plain loop: {t_plain * 1e3:.2f} ms")
    batch.close()
    # the same comparison one grid size at a time
    per_size = {}
    for N in BATCH_SIZES:
        sub = [b for b in boxes if b["params"].grid_dim == N]
        bs = reion.HiiBatch(ctx, sub, graph=True, tune=True)
        bs.run(sync=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            bs.run(sync=False)
        torch.cuda.synchronize()
        tb = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        for b in bs.boxes:
            reion.find_hii_bubbles(ctx, b["params"], b["deltax"], b["stars"], b["weighted_sfr"],
                                   z_at_ionization=b["z_at_ionization"], out=b)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        bs.close()
        alg = sum(path_bytes_hii(N, 3, len(P.radius_ladder(b["params"]))) for b in sub)
        per_size[str(N)] = {"batched_ms": tb * 1e3, "plain_loop_ms": tp * 1e3, "speedup": tp / tb,
                            "alg_bytes": alg, "batched_GB_s": alg / tb / 1e9, "batched_frac": alg / tb / HBM_PEAK,
                            "plain_loop_GB_s": alg / tp / 1e9, "plain_loop_frac": alg / tp / HBM_PEAK}
        log(f"[bench] batched {N}^3 x {len(sub)}: {tb * 1e3:.2f} ms ({alg / tb / 1e9:.0f} GB/s, "
            f"{alg / tb / HBM_PEAK:.2f} of peak), plain loop {tp * 1e3:.2f} ms ({alg / tp / 1e9:.0f} GB/s, "
            f"{alg / tp / HBM_PEAK:.2f}), x{tp / tb:.2f}")
    return {
        "metric": "cells*radii/s of find_HII_bubbles over 96 uneven boxes, batched (one graph launch)",
        "value": cells_radii / t_batch,
        "unit": "cells*radii/s",
        "ms_per_launch": t_batch * 1e3,
        "plain_loop_value": cells_radii / t_plain,
        "plain_loop_ms": t_plain * 1e3,
        "speedup_vs_plain_loop": t_plain / t_batch,
        "model_hbm_frac": mbytes / t_batch / HBM_PEAK,
        "model_bytes": mbytes,
        "cells_radii": cells_radii,
        "plan_s": t_plan,
        "per_grid": per_size,
        "graph_nodes": info["graph_nodes"],
        "cohorts": [{k: c[k] for k in ("grid_dim", "boxes", "radii", "variant", "tile_width", "radix_schedule")}
                    for c in info["cohorts"]],
        "config": {"workload": "32 x 64^3 + 32 x 128^3 + 32 x 192^3 boxes, 100 Mpc/h, z=8, top-hat, 3 fields, "
                               "eps U[10,40], synthetic (torch generator)"},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["snapshot", "bubbles", "batched", "gridding", "grid_files"],
                    default="snapshot")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the adjoint, batched and gridding parts of the snapshot run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 3 times (scale of the PMC byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the reionization path runs the whole box on a single MI355X "
              f"(--gpus {args.gpus} is not supported; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from meraxes_amd import params as P
    from meraxes_amd import reion, synth
    from meraxes_amd._lib import Context
    from meraxes_amd.snapshots import SyntheticRun

    if args.workload == "batched":
        ctx = Context(0)
        b = bench_batched(ctx, args.steps, args.warmup)
        ctx.close()
        out = {"metric": b["metric"], "value": b["value"], "unit": b["unit"], "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": b["ms_per_launch"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": b["config"],
               "batched": b}
        print(json.dumps(out), flush=True)
        return

    if args.workload == "gridding":
        ctx = Context(0)
        b = bench_gridding(ctx, N=args.grid, reps=max(args.steps, 1))
        ctx.close()
        out = {"metric": b["metric"], "value": b["value"], "unit": "galaxies/s", "n_gpus": 1, "steps": args.steps,
               "warmup": 1, "ms_per_step": b["ms_per_call"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": {"workload": b["config"]},
               "gridding": b}
        print(json.dumps(out), flush=True)
        return

    if args.workload == "grid_files":
        ctx = Context(0)
        b = bench_grid_files(ctx, N=args.grid, reps=max(args.steps, 1))
        ctx.close()
        out = {"metric": "bytes/s of one snapshot's HDF5 grid files written from HBM", "value": b["GB_s"] * 1e9,
               "unit": "bytes/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
               "ms_per_step": b["ms_per_snapshot"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": {"workload": b["config"]},
               "grid_files": b}
        print(json.dumps(out), flush=True)
        return

    N = args.grid
    L = 400.0 * N / 512
    snapshot = args.workload == "snapshot"
    zz = P.snapshot_redshifts()
    snap = 24  # z = 8.41 in the 30-snapshot run from z = 20 to 6
    z = float(zz[snap])
    log(f"[bench] generating {N}^3 inputs ({args.workload})")
    # the full-snapshot configuration's density: one seeded k^-2 field grown
    # linearly as 1/(1+z) from z = 20, sources from the grown field
    d, s, w = SyntheticRun(N, L, seed=0).inputs(snap)
    dev = torch.device("cuda:0")
    dd, ds, dw = (torch.from_numpy(a).to(dev) for a in (d, s, w))
    del d, s, w
    F = 4 if snapshot else 3
    p = P.hii_params(N, L, z, filter_type=0, efficiency=20.0, uvb_flag=1, include_spin_temp=1 if snapshot else 0)
    tp = P.tb_params(N, z, include_spin_temp=1 if snapshot else 0)
    S = len(P.radius_ladder(p))
    xH = torch.empty((N, N, N), dtype=torch.float32, device=dev)
    rb = torch.empty_like(xH)
    J21 = torch.empty_like(xH)
    dT = torch.empty_like(xH)
    zin = torch.full_like(xH, -1.0)
    J21i = torch.zeros_like(xH)
    ctx = Context(0)
    if snapshot:
        # carried state and SFR history: earlier snapshots' SFR grids scaled by the
        # linear growth of the source field, (1+z)^-3 relative to this one
        nh = 8
        hist = torch.empty((nh,) + tuple(dw.shape), dtype=torch.float32, device=dev)
        for k in range(nh):
            zk = zz[max(snap - k, 0)]
            hist[k] = dw * float(((1 + z) / (1 + zk)) ** 3)
        fcoll = np.zeros(len(zz))
        fcoll[:snap] = np.linspace(1e-4, 5e-3, snap)
        tsp = P.ts_params(N, L, snap, zz, fcoll, nh, efficiency=20.0)
        xe0 = torch.full_like(dd, 2e-3)
        Tk0 = torch.full_like(xH, 50.0)
        xe_prev = torch.empty_like(dd)
        Tk = torch.empty_like(xH)
        xe_box = torch.empty_like(dd)
        TS = torch.empty_like(xH)

    def step():
        zin.fill_(-1.0)
        xe_in = None
        TSb = None
        if snapshot:
            xe_prev.copy_(xe0)
            Tk.copy_(Tk0)
            reion.compute_ts(ctx, tsp, dd, ds, dw, hist, xe_prev, Tk, x_e_box=xe_box, TS_box=TS)
            xe_in = xe_box
            TSb = TS
        r = reion.find_hii_bubbles(ctx, p, dd, ds, dw, x_e=xe_in, z_at_ionization=zin, J_21_at_ionization=J21i,
                                   out={"xH": xH, "r_bubble": rb, "J_21": J21})
        reion.brightness_temp(ctx, tp, xH, dd, TSb, out=dT)
        return r

    # the steps run on a torch side stream (default priority), which the
    # library enqueues on directly: on torch's legacy default stream every
    # call would first wait for the host to see the device idle
    side = torch.cuda.Stream()
    log(f"[bench] warmup {args.warmup}")
    with torch.cuda.stream(side):
        for _ in range(args.warmup):
            r = step()
    torch.cuda.synchronize()
    reion.profile(ctx, True)
    log(f"[bench] timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(side):
        for _ in range(args.steps):
            r = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    prof = reion.profile_read(ctx)
    reion.profile(ctx, False)
    log(f"[bench] <xH>_V = {r['volume_weighted_global_xH']:.6f}, {dt * 1e3:.2f} ms/step")
    for k, (ms, n) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        log(f"[bench]   {k:18s} {ms / args.steps:9.3f} ms/step  {n // args.steps:5d} launches/step  "
            f"{ms / n * 1e3:9.1f} us/launch")

    # dominant kernel roofline from the event timings
    dom = max(prof.items(), key=lambda kv: kv[1][0])
    name, (ms_tot, cnt) = dom
    avg_s = ms_tot / cnt / 1e3
    # radii per multi-radius x pass: the S - 1 filtered radii over its launches
    nk = prof.get("ix_filter_k", (0, 0))[1] // args.steps
    radii = (S - 1) / nk if nk else 1.0
    b = kernel_bytes(name, N, F, radii)
    roof = {"kernel": name, "bound": "hbm", "achieved": (b / avg_s / 1e9) if b else None, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": (b / avg_s / HBM_PEAK) if b else None, "traffic": measured_traffic(name, N),
            "avg_launch_us": avg_s * 1e6, "bytes_per_launch": b}
    per_kernel = {}
    for k, (ms, n) in prof.items():
        kb = kernel_bytes(k, N, F, radii)
        per_kernel[k] = {"ms_per_step": ms / args.steps, "launches_per_step": n // args.steps,
                         "us_per_launch": ms / n * 1e3,
                         "GB_s": (kb * n / (ms / 1e3) / 1e9) if kb else None,
                         "frac": (kb * n / (ms / 1e3) / HBM_PEAK) if kb else None}
    Bc = 8 * N * N * (N // 2 + 1)
    mb_hii = model_bytes_hii(N, F, S, True)
    mb = mb_hii + 16 * N**3
    if snapshot:
        mb += 40 * (12 * Bc + 8 * N**3) + 184 * N**3  # the 3-pass model of the Ts shells and per-cell pass
    # device time per stage (sums of the kernels' event timings)
    stage = {"ts": sum(v[0] for k, v in prof.items() if k.startswith("ts_")) / args.steps,
             "dTb": sum(v[0] for k, v in prof.items() if k == "brightness_temp") / args.steps}
    stage["bubbles"] = sum(v[0] for v in prof.values()) / args.steps - stage["ts"] - stage["dTb"]
    cells_radii = N**3 * S
    out = {
        "metric": ("cells*radii/s of one snapshot (Ts + find_HII_bubbles + dTb) at 512^3, plain path" if snapshot
                   else "cells*radii/s of find_HII_bubbles + dTb at 512^3, plain path"),
        "value": cells_radii / dt,
        "unit": "cells*radii/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{'ComputeTs (40 shells) + ' if snapshot else ''}find_HII_bubbles + dTb, "
                                f"{N}^3 grid, {L:g} Mpc/h, z={z:.3f}, top-hat, UVB on, eps=20, {S} radii, "
                                f"{F} fields"), "grid": N, "box_mpc_h": L, "radii": S,
                   "fields": F, "ts_shells": 40 if snapshot else 0, "parallelism": "single GPU"},
        "model_hbm_frac": mb / dt / HBM_PEAK,
        "model_bytes": mb,
        "model_note": ("bytes of the 3-pass model of the reference's structure (separate filter, clamp and "
                       "threshold traffic, float state grids, dense shell transforms) over the step's wall time; "
                       "this path fuses the filter, clamp and threshold into the transform passes, keeps one byte "
                       "of ladder state per cell and touches only the live modes of the sharp-k shells, so it "
                       "moves fewer bytes than the model: the kernels' GB_s (their own algorithmic bytes) and "
                       "roofline.traffic (PMC) are the HBM figures"),
        "bubbles_model_frac": mb_hii / (stage["bubbles"] / 1e3) / HBM_PEAK,
        "stage_device_ms": stage,
        "roofline": roof,
        "kernels": per_kernel,
    }
    # the path at the kernels' own algorithmic bytes: every launch whose bytes are
    # known (the sharp-k shells' pruned passes are not counted) over the step time
    known = sum((kernel_bytes(k, N, F, radii) or 0) * n for k, (ms, n) in prof.items()) / args.steps
    known_ms = sum(ms for k, (ms, n) in prof.items() if kernel_bytes(k, N, F, radii)) / args.steps
    out["path_kernel_bytes"] = {
        "bytes_per_step": known, "frac_of_peak": known / dt / HBM_PEAK,
        "device_ms_of_those_kernels": known_ms, "frac_of_peak_over_their_device_time": known / (known_ms / 1e3) / HBM_PEAK,
        "note": "sum over launches of each kernel's algorithmic bytes (kernels[].GB_s), over the step's wall time "
                "and over those kernels' own device time; the pruned sharp-k shell passes are left out"}
    out["hbm_copy_ceiling"] = copy_ceiling(torch, ctx)
    log(f"[bench] device copy ceiling {out['hbm_copy_ceiling']['GB_s']:.0f} GB/s streaming kernel, "
        f"{out['hbm_copy_ceiling']['torch_copy_GB_s']:.0f} GB/s torch copy_ (read + write)")
    if snapshot and not args.no_extras:
        del dd, ds, dw, hist, xH, rb, J21, dT, zin, J21i, xe0, Tk0, xe_prev, Tk, xe_box, TS
        torch.cuda.empty_cache()
        out["adjoint"] = {str(n): bench_adjoint(ctx, n) for n in (256, N)}
        out["batched"] = bench_batched(ctx, 2, 1)
        out["gridding"] = bench_gridding(ctx)
        out["tb_velocities"] = bench_tb_vel(ctx)
        out["grid_files"] = bench_grid_files(ctx)
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(snapshot)
        log(f"[bench] CPU baseline {out['cpu_baseline']['value']:.3e} cells*radii/s "
            f"on {out['cpu_baseline']['cores']} threads")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
