"""Benchmark of the pixel-space beam convolution on one MI355X.

Default workload (the headline metric): forward + adjoint (TOD -> I,Q,U map)
of a detector pair at sky NSIDE 512 with a 5 degree beam cap:
  sky    NSIDE 512 RING, float32 I,Q,U power-law Gaussian random fields
         (C_l ~ l^-2, l <= 512), seeded
  beams  elliptical-Gaussian pair, FWHM 0.33 deg, cap 5 deg, beam NSIDE 4096
         (382,812 beam pixels per detector)
  scan   4e6 samples of a constant-elevation whole-sky scan (lat -23 deg,
         el 45 deg, 1 rpm, 100 Hz, sidereal drift; dec in [-68, +22] deg)
A step is one forward over all samples followed by one adjoint-map over all
samples (cotangents = a fixed random vector), inputs resident in HBM.

Roofline (gather model): a visit is one (sample, disc pixel) pair and reads
the sky I,Q,U (12 B) and both beams' I,Q,U (24 B); per sample 12 B of
pointing plus 4 B per detector of TOD (written by the forward, read as the
cotangent by the adjoint); the adjoint also reads and writes the 12 B I,Q,U
gradient of every visit:
  B_fwd  = sum_i |D_i| * 36 + N * 20,   B_adjm = sum_i |D_i| * 48 + N * 20
with |D_i| the exact disc sizes (device query_disc).  achieved = bytes of
the dominant kernel / its mean duration (HIP events on the launch stream).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--samples N]
                  [--no-cpu] [--no-forms] [--no-batched] [--no-mapmaker] [--no-c5]
                  [--beam-adjoint-samples M] [--calib-copy-mib M]

The beam adjoint (TOD -> beam maps) is timed over the headline scan (or the
prefix --beam-adjoint-samples), with B_adjb = sum_i |D_i| * 60 + N * 20; the
annulus-band form it replaced is timed beside it on the first 2.5e5 samples.

Also reported: the plain drop-in forward of the NSIDE 256 / 1.5 deg workload
(device-resident and through the host API including PCIe copies), the
batched 64-detector forward + adjoint against the plain per-pair launches,
the binned map-maker on 64 detectors, and the CPU oracle on a prefix of the
headline workload.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12   # bytes/s, MI355X
# HBM-side bytes per launch of the headline workload's kernels, measured with
# rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a copy of known size)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_nside512_cap5_4e6.txt")
# hardware counters (VALU and TA busy, waits, L2 hits) of the same kernels on a
# 1e6-sample prefix, measured with rocprofv3 --pmc
PMC_FILE = os.path.join(ROOT, "profiles", "pmc_kernels_nside512_cap5_1e6.txt")


def measured_limiter(kernel):
    """VALU busy, TA busy, wait fraction and L2 hit rate of `kernel` (its
    largest grid) from the committed counter summary, and what they say
    bounds it; None if the file or the kernel is absent."""
    try:
        lines = open(PMC_FILE).read().splitlines()
    except OSError:
        return None
    best = None
    for l in lines:
        t = l[30:].split()
        if len(t) != 9 or l[:30].strip() != kernel:
            continue
        try:
            grid, valu, ta, wait, hit = int(t[-9]), float(t[-7]), float(t[-6]), float(t[-5]), \
                float(t[-4])
        except ValueError:
            continue
        if best is None or grid > best[0]:
            best = (grid, valu, ta, wait, hit)
    if best is None:
        return None
    grid, valu, ta, wait, hit = best
    if valu >= 0.95:
        lim = "VALU issue"
    elif ta >= 0.85:
        lim = "texture addressing (TA) of the gathers"
    else:
        lim = "latency: VALU %.0f%% busy, waits %.0f%% of wave cycles" % (100 * valu, 100 * wait)
    return {"valu_busy": valu, "ta_busy": ta, "wait_frac": wait, "l2_hit_rate": hit,
            "limiter": lim, "counters": os.path.relpath(PMC_FILE, ROOT)}


def measured_traffic(samples, kernel, what="traffic_bytes_per_launch"):
    """Calibrated HBM bytes per launch of `kernel` (or, with what =
    "l2_hit_rate", its L2 hit rate) from the committed profile, if it was
    measured on this sample count; else None."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    vals = dict(l.split(": ", 1) for l in lines if ": " in l and not l.startswith(" "))
    if vals.get("samples") != str(samples):
        return None
    v = vals.get("%s %s" % (what, kernel))
    return float(v) if v else None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--samples", type=int, default=4_000_000)
    p.add_argument("--cpu-samples", type=int, default=20_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-forms", action="store_true",
                   help="skip the plain drop-in forward of the NSIDE 256 workload")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the batched 64-detector workload")
    p.add_argument("--no-c5", action="store_true",
                   help="skip the large configuration (NSIDE 2048, 256 detectors)")
    p.add_argument("--beam-adjoint-samples", type=int, default=-1,
                   help="prefix of the headline scan on which the beam adjoint is timed "
                        "(-1: the whole scan; 0: skip)")
    p.add_argument("--no-mapmaker", action="store_true",
                   help="skip the binned map-maker workload")
    p.add_argument("--calib-copy-mib", type=int, default=0,
                   help="run one device-to-device copy of this size first (a kernel of "
                        "known traffic, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from pisco4cmb_amd import inputs, torch_ops
    from pisco4cmb_amd._lib import check, lib

    dev = torch.device("cuda:0")
    if args.calib_copy_mib:
        src = torch.ones(args.calib_copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    nside_sky, nside_beam, cap_deg, fwhm = 512, 4096, 5.0, 0.33
    n = args.samples

    t0 = time.time()
    I, Q, U, _ = inputs.powerlaw_sky(nside_sky, seed=2024, lmax=512)
    beam = inputs.gaussian_polbeam(nside_beam, cap_deg, fwhm, 0.0, 0.0, nside_sky=nside_sky)
    ra, dec, pa = inputs.ces_scan(n)
    g = np.random.default_rng(7).normal(size=(2, n)).astype(np.float32)
    setup_s = time.time() - t0

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    It, Qt, Ut = t(I), t(Q), t(U)
    ba, bb = t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams))
    rat, dect, pat = t(ra), t(dec), t(pa)
    gat, gbt = t(g[0]), t(g[1])
    beam8 = torch_ops.pack_beam(ba, bb)
    npb = beam8.shape[0]
    npix_sky = 12 * nside_sky * nside_sky
    tod_a = torch.empty(n, dtype=torch.float32, device=dev)
    tod_b = torch.empty_like(tod_a)
    gI = torch.empty(npix_sky, dtype=torch.float32, device=dev)
    gQ = torch.empty_like(gI)
    gU = torch.empty_like(gI)
    ctx = torch_ops._ctx(dev)
    stream = torch_ops._stream(dev)

    # exact disc sizes (device query_disc, the same code the kernels run)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    rho_max = lib.pisco_beam_rho_max(nside_beam, npb)
    check(lib.pisco_dev_query_disc(ctx, stream, nside_sky, n, rat.data_ptr(), dect.data_ptr(),
                                   rho_max, None, counts.data_ptr(), None))
    torch.cuda.synchronize()
    visits = int(counts.to(torch.int64).sum().item())
    k = 2
    b_fwd = visits * (12 + 12 * k) + n * (12 + 4 * k)
    b_adj = visits * (12 * k + 24) + n * (12 + 4 * k)

    def fwd():
        check(lib.pisco_dev_forward(ctx, stream, nside_sky, It.data_ptr(), Qt.data_ptr(),
                                    Ut.data_ptr(), nside_beam, npb, beam8.data_ptr(), n,
                                    rat.data_ptr(), dect.data_ptr(), pat.data_ptr(),
                                    tod_a.data_ptr(), tod_b.data_ptr()))

    def adj():
        check(lib.pisco_dev_adjoint_map(ctx, stream, nside_sky, nside_beam, npb,
                                        beam8.data_ptr(), n, rat.data_ptr(), dect.data_ptr(),
                                        pat.data_ptr(), gat.data_ptr(), gbt.data_ptr(),
                                        gI.data_ptr(), gQ.data_ptr(), gU.data_ptr()))

    for _ in range(args.warmup):
        fwd()
        adj()
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.time()
    start.record()
    for e0, e1, e2 in ev:
        e0.record()
        fwd()
        e1.record()
        adj()
        e2.record()
    stop.record()
    torch.cuda.synchronize()
    wall = time.time() - wall0
    total_ms = start.elapsed_time(stop)
    t_fwd = sum(a.elapsed_time(b) for a, b, _ in ev) / args.steps / 1e3
    t_adj = sum(b.elapsed_time(c) for _, b, c in ev) / args.steps / 1e3
    ms_per_step = total_ms / args.steps
    value = k * n / (ms_per_step / 1e3)

    tod_host = (tod_a[:args.cpu_samples].cpu().numpy(), tod_b[:args.cpu_samples].cpu().numpy())

    # the sky adjoint with float sums (the timed steps use the default integer
    # sums, which are deterministic), outside the timed steps
    check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", 0))
    adj()
    d0 = torch.cuda.Event(enable_timing=True)
    d1 = torch.cuda.Event(enable_timing=True)
    d0.record()
    adj()
    d1.record()
    torch.cuda.synchronize()
    t_adj_float = d0.elapsed_time(d1) / 1e3
    check(lib.pisco_set_tuning(ctx, b"adjmap_deterministic", -1))

    # the beam adjoint (TOD -> beam maps) on a prefix, outside the timed steps
    nbp = n if args.beam_adjoint_samples < 0 else min(n, args.beam_adjoint_samples)
    beam_adj = None
    if nbp > 0:
        gbeam8 = torch.empty(npb * 8, dtype=torch.float32, device=dev)

        def adjb(m):
            check(lib.pisco_dev_adjoint_beam(ctx, stream, nside_sky, It.data_ptr(), Qt.data_ptr(),
                                             Ut.data_ptr(), nside_beam, npb, m, rat.data_ptr(),
                                             dect.data_ptr(), pat.data_ptr(), gat.data_ptr(),
                                             gbt.data_ptr(), gbeam8.data_ptr()))

        def time_adjb(m):
            adjb(m)
            d0.record()
            adjb(m)
            d1.record()
            torch.cuda.synchronize()
            return d0.elapsed_time(d1) / 1e3

        def model(m):
            v_pre = int(counts[:m].to(torch.int64).sum().item())
            return v_pre * (12 + 24 * k) + m * (12 + 4 * k)
        t_adjb = time_adjb(nbp)
        b_adjb = model(nbp)
        beam_adj = {"samples": nbp, "s": t_adjb, "tod_samples_per_s": k * nbp / t_adjb,
                    "hbm_frac": b_adjb / t_adjb / HBM_PEAK,
                    "form": "visit records in band order (two mapping passes), band sums in "
                            "LDS as 64-bit fixed point",
                    "measured_limiters": {k: measured_limiter(k) for k in
                                          ("k_beamrec<false>", "k_beamrec<true>",
                                           "k_beamrec_sum")}}
        # the annulus-band form, on a prefix
        nba = min(nbp, 250_000)
        check(lib.pisco_set_tuning(ctx, b"adjbeam_records", 0))
        t_band = time_adjb(nba)
        check(lib.pisco_set_tuning(ctx, b"adjbeam_records", 1))
        t_rec = time_adjb(nba)
        beam_adj["annulus_bands"] = {"samples": nba, "s": t_band,
                                     "hbm_frac": model(nba) / t_band / HBM_PEAK,
                                     "records_s_same_prefix": t_rec}
        del gbeam8

    dominant = "adjoint_map" if t_adj >= t_fwd else "forward"
    b_dom, t_dom = (b_adj, t_adj) if dominant == "adjoint_map" else (b_fwd, t_fwd)
    k_dom = "k_adjmap_lds" if dominant == "adjoint_map" else "k_pair<0>"
    # the timed sky adjoint is the integer-sum (deterministic) form
    k_traffic = "k_adjmap_lds_deterministic" if k_dom == "k_adjmap_lds" else k_dom
    achieved = b_dom / t_dom
    out = {
        "metric": "TOD samples/s fwd+adjoint @NSIDE512 5deg beam (pair); %HBM peak",
        "value": value,
        "unit": "TOD samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 data/products, f64 index path + accumulation",
        "data": "synthetic (seeded power-law sky, Gaussian beams, CES scan)",
        "config": {"workload": "fwd+adjoint-map, sky NSIDE 512, pair, FWHM 0.33 deg, cap "
                               "5 deg, beam NSIDE 4096, %d samples whole-sky CES" % n,
                   "nside_sky": nside_sky, "nside_beam": nside_beam, "npix_beam": npb,
                   "samples": n, "detectors": k, "visits_per_pass": visits,
                   "mean_disc": visits / n},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(n, k_traffic),
                     "l2_hit_rate": measured_traffic(n, k_traffic, "l2_hit_rate"),
                     "kernel": k_dom, "bytes_per_launch": b_dom, "mean_launch_s": t_dom,
                     "measured_limiter": measured_limiter(
                         "k_adjmap_lds<0, true, 64>" if k_dom == "k_adjmap_lds" else k_dom),
                     "note": "frac prices the gather model against HBM; the counters show "
                             "what bounds the kernel (measured_limiter)"},
        "fwd_adj": {"forward_s": t_fwd, "adjoint_s": t_adj,
                    "adjoint_deterministic": True,
                    "adjoint_float_sums_s": t_adj_float,
                    "frac_hbm_fwd_plus_adj": (b_fwd + b_adj) / (t_fwd + t_adj) / HBM_PEAK,
                    "forward_tod_samples_per_s": k * n / t_fwd,
                    "forward_measured_limiter": measured_limiter("k_pair<0>"),
                    "wall_s_timed": wall, "setup_s": setup_s},
    }
    if beam_adj:
        out["beam_adjoint"] = beam_adj

    if not args.no_forms:
        out["plain_forward_nside256"] = plain_forward(ctx, stream, dev)
    if not args.no_batched:
        out["batched_64det"] = batched_workload(ctx, stream, dev, It, Qt, Ut, nside_sky)
        out["batched_5deg_16det"] = batched_workload(ctx, stream, dev, It, Qt, Ut, nside_sky,
                                                     npairs=8, five_deg=True)
    if not args.no_c5:
        out["large_256det"] = large_workload(dev)
    if not args.no_mapmaker:
        out["map_maker_64det"] = map_maker_workload(dev, nside_sky, with_cpu=not args.no_cpu)
    if not args.no_cpu:
        cb = cpu_baseline(args, I, Q, U, nside_sky, nside_beam, beam, ra, dec, pa, g, tod_host)
        parity = cb.pop("parity", None)
        out["cpu_baseline"] = cb
        if parity:
            out["parity"] = parity
    print(json.dumps(out), flush=True)


def plain_forward(ctx, stream, dev, n=1_000_000, reps=5):
    """The plain drop-in forward of a detector pair: sky NSIDE 256 (power-law,
    l <= 256), beam NSIDE 2048, cap 1.5 deg, FWHM 0.5 deg, 1e6 CES samples.
    Timed two ways: device-resident (inputs in HBM, HIP events) and through
    the host API the reference exposes (GPUConvolver.update_scan +
    exec_convolution with host arrays: includes the PCIe copies)."""
    import numpy as np
    import torch

    from pisco4cmb_amd import inputs, pypisco, torch_ops
    from pisco4cmb_amd._lib import check, lib
    nside_sky, nside_beam = 256, 2048
    I, Q, U, _ = inputs.powerlaw_sky(nside_sky, seed=77, lmax=256)
    beam = inputs.gaussian_polbeam(nside_beam, 1.5, 0.5, 0.0, 0.0, nside_sky=nside_sky)
    ra, dec, pa = inputs.ces_scan(n)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    It, Qt, Ut, rat, dect, pat = (t(x) for x in (I, Q, U, ra, dec, pa))
    beam8 = torch_ops.pack_beam(t(np.stack(beam.aBeams)), t(np.stack(beam.bBeams)))
    npb = beam8.shape[0]
    ta = torch.empty(n, dtype=torch.float32, device=dev)
    tb = torch.empty_like(ta)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    check(lib.pisco_dev_query_disc(ctx, stream, nside_sky, n, rat.data_ptr(), dect.data_ptr(),
                                   lib.pisco_beam_rho_max(nside_beam, npb), None,
                                   counts.data_ptr(), None))
    torch.cuda.synchronize()
    visits = int(counts.to(torch.int64).sum().item())

    def fwd():
        check(lib.pisco_dev_forward(ctx, stream, nside_sky, It.data_ptr(), Qt.data_ptr(),
                                    Ut.data_ptr(), nside_beam, npb, beam8.data_ptr(), n,
                                    rat.data_ptr(), dect.data_ptr(), pat.data_ptr(),
                                    ta.data_ptr(), tb.data_ptr()))
    gat, gbt = (t(x) for x in np.random.default_rng(78).normal(size=(2, n)).astype(np.float32))
    gI, gQ, gU = (torch.empty(12 * nside_sky ** 2, dtype=torch.float32, device=dev)
                  for _ in range(3))

    def adj():
        check(lib.pisco_dev_adjoint_map(ctx, stream, nside_sky, nside_beam, npb, beam8.data_ptr(),
                                        n, rat.data_ptr(), dect.data_ptr(), pat.data_ptr(),
                                        gat.data_ptr(), gbt.data_ptr(), gI.data_ptr(),
                                        gQ.data_ptr(), gU.data_ptr()))

    def timed(f):
        f()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3
    t_dev = timed(fwd)
    t_adj = timed(adj)
    # the same calls with one sample per wave (the kernels the large discs use)
    check(lib.pisco_set_tuning(ctx, b"forward_lanes", 64))
    check(lib.pisco_set_tuning(ctx, b"adjmap_lanes", 64))
    t_dev64, t_adj64 = timed(fwd), timed(adj)
    check(lib.pisco_set_tuning(ctx, b"forward_lanes", 0))
    check(lib.pisco_set_tuning(ctx, b"adjmap_lanes", 0))
    b = visits * 36 + n * 20
    b_adj = visits * 48 + n * 20

    conv = pypisco.GPUConvolver(pypisco.default_run_config(ptgPerConv=n))
    conv.configure_execution()
    conv.update_sky(pypisco.Sky(nside_sky, I, Q, U))
    conv.update_beam(beam)
    da = np.zeros(n, np.float32)
    db = np.zeros(n, np.float32)
    scan = pypisco.Scan(n, ra, dec, pa)
    conv.update_scan(scan)
    conv.exec_convolution(da, db)
    w0 = time.time()
    for _ in range(reps):
        conv.update_scan(scan)
        conv.exec_convolution(da, db)
    t_host = (time.time() - w0) / reps
    conv.close()
    return {"workload": "forward, sky NSIDE 256, pair, FWHM 0.5 deg, cap 1.5 deg, beam "
                        "NSIDE 2048, %d samples whole-sky CES" % n,
            "visits": visits, "mean_disc": visits / n,
            "device_resident": {"tod_samples_per_s": 2 * n / t_dev, "s": t_dev,
                                "hbm_frac": b / t_dev / HBM_PEAK},
            "host_api_with_pcie": {"tod_samples_per_s": 2 * n / t_host, "s": t_host},
            "adjoint_device_resident": {"tod_samples_per_s": 2 * n / t_adj, "s": t_adj,
                                        "hbm_frac": b_adj / t_adj / HBM_PEAK},
            "fwd_adj_device_resident": {"tod_samples_per_s": 2 * n / (t_dev + t_adj),
                                        "hbm_frac": (b + b_adj) / (t_dev + t_adj) / HBM_PEAK},
            "one_sample_per_wave": {"forward_s": t_dev64, "adjoint_s": t_adj64}}


def batched_workload(ctx, stream, dev, It, Qt, Ut, nside_sky, n=250_000, npairs=32, reps=3,
                     five_deg=False):
    """64 detectors (32 pairs), each pair with its own elliptical Gaussian beam
    (FWHM U[0.2, 1.0] deg, ellipticity U[0, 0.1], cap 3 x FWHM, beam NSIDE
    2048) and position-angle offset, on the headline sky and a 2.5e5-sample
    CES scan: the batched forward + sky adjoint (one launch each, the step
    replayed from a hipGraph) against the plain form (one single-pair launch
    per pair and operation).  five_deg: the pairs have the headline's beam
    shape instead (cap 5 deg, beam NSIDE 4096, FWHM U[0.30, 0.36] deg,
    ellipticity U[0, 0.1])."""
    import numpy as np
    import torch

    from pisco4cmb_amd import inputs, torch_ops
    from pisco4cmb_amd._lib import check, lib
    from pisco4cmb_amd.batched import BatchedStepGraph, DetectorBatch
    rng = np.random.default_rng(65 if five_deg else 64)
    ra, dec, pa = inputs.ces_scan(n, t0=3600.0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rat, dect, pat = t(ra), t(dec), t(pa)
    specs, radii = [], []
    for _ in range(npairs):
        if five_deg:
            fwhm, nb = float(rng.uniform(0.30, 0.36)), 4096
            cap = 5.0
        else:
            fwhm, nb = float(rng.uniform(0.2, 1.0)), 2048
            cap = 3 * fwhm
        b = inputs.gaussian_polbeam(nb, cap, fwhm, float(rng.uniform(0, 0.1)), 0.0,
                                    nside_sky=nside_sky)
        specs.append((t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb,
                      float(rng.uniform(-np.pi, np.pi))))
        radii.append(b.get_rho_max())
    batch = DetectorBatch(specs)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    visits = 0
    for r in radii:
        check(lib.pisco_dev_query_disc(ctx, stream, nside_sky, n, rat.data_ptr(),
                                       dect.data_ptr(), r, None, counts.data_ptr(), None))
        visits += int(counts.to(torch.int64).sum().item())
    g = torch.randn((npairs, 2, n), device=dev)

    def tm(f):
        f()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_bf = tm(lambda: batch.forward(It, Qt, Ut, rat, dect, pat, nside_sky))
    t_ba = tm(lambda: batch.adjoint_map(g, rat, dect, pat, nside_sky))
    step = BatchedStepGraph(batch, It, Qt, Ut, rat, dect, pat, g, nside_sky)
    t_step = tm(step.replay)
    pas = [t((pa + np.float32(sp[3])).astype(np.float32)) for sp in specs]
    b8 = [torch_ops.pack_beam(sp[0], sp[1]) for sp in specs]
    ta = torch.empty(n, device=dev)
    tb = torch.empty(n, device=dev)
    gI = torch.empty(12 * nside_sky ** 2, device=dev)
    gQ = torch.empty_like(gI)
    gU = torch.empty_like(gI)

    def plain_fwd():
        for k in range(npairs):
            check(lib.pisco_dev_forward(ctx, stream, nside_sky, It.data_ptr(), Qt.data_ptr(),
                                        Ut.data_ptr(), specs[k][2], b8[k].shape[0],
                                        b8[k].data_ptr(),
                                        n, rat.data_ptr(), dect.data_ptr(), pas[k].data_ptr(),
                                        ta.data_ptr(), tb.data_ptr()))

    def plain_adj():
        for k in range(npairs):
            check(lib.pisco_dev_adjoint_map(ctx, stream, nside_sky, specs[k][2], b8[k].shape[0],
                                            b8[k].data_ptr(), n, rat.data_ptr(),
                                            dect.data_ptr(), pas[k].data_ptr(),
                                            g[k, 0].data_ptr(), g[k, 1].data_ptr(),
                                            gI.data_ptr(), gQ.data_ptr(), gU.data_ptr()))

    t_pf = tm(plain_fwd)
    t_pa = tm(plain_adj)
    batch.close()
    det = 2 * npairs * n
    b_fa = visits * (36 + 48) + 2 * npairs * n * 20
    shape = ("FWHM U[0.30,0.36] deg, cap 5 deg, beam NSIDE 4096" if five_deg
             else "FWHM U[0.2,1.0] deg, cap 3xFWHM, beam NSIDE 2048")
    return {"workload": "%d detectors (%d pairs), %s, sky NSIDE %d, %d samples CES, per-pair "
                        "pa offsets" % (2 * npairs, npairs, shape, nside_sky, n),
            "pair_visits": visits,
            "batched_fwd_adj_graph": {"det_samples_per_s": det / t_step, "s": t_step,
                                      "hbm_frac": b_fa / t_step / HBM_PEAK},
            "batched_forward": {"det_samples_per_s": det / t_bf, "s": t_bf},
            "batched_adjoint": {"det_samples_per_s": det / t_ba, "s": t_ba},
            "plain_forward": {"det_samples_per_s": det / t_pf, "s": t_pf},
            "plain_fwd_adj": {"det_samples_per_s": det / (t_pf + t_pa), "s": t_pf + t_pa,
                              "hbm_frac": b_fa / (t_pf + t_pa) / HBM_PEAK},
            "batched_fwd_adj_vs_plain_forward": t_pf / t_step,
            "batched_fwd_adj_vs_plain_fwd_adj": (t_pf + t_pa) / t_step,
            "measured_limiters": {k: measured_limiter(k) for k in
                                  ("k_batch_fwd<4>", "k_adjmap_lds<4, true, 64>")}}


def large_workload(dev, n=100_000, per_family=32, reps=1):
    """The large configuration: sky NSIDE 2048 (power-law I, Q, U, l <= 128,
    plus 200 point sources), 256 detectors as 128 pairs in four beam families
    (FWHM 0.1, 0.2, 0.5, 1.0 deg; cap 3 x FWHM; beam NSIDE 8192, 8192, 4096,
    2048; ellipticity U[0, 0.1]; per-pair position-angle offsets), 1e5 CES
    samples shared by the pairs.  One batched forward and one batched sky
    adjoint, timed with HIP events; gather-model bytes as the other legs."""
    import numpy as np
    import torch

    from pisco4cmb_amd import inputs
    from pisco4cmb_amd._lib import check, lib
    from pisco4cmb_amd.batched import DetectorBatch
    from pisco4cmb_amd import torch_ops
    ns = 2048
    I, Q, U, _ = inputs.powerlaw_sky(ns, seed=2048, lmax=128)
    inputs.add_point_sources(I, ns, seed=2049)
    ra, dec, pa = inputs.ces_scan(n, t0=1800.0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    It, Qt, Ut, rat, dect, pat = (t(x) for x in (I, Q, U, ra, dec, pa))
    rng = np.random.default_rng(2050)
    specs, radii = [], []
    for fwhm, nb in ((0.1, 8192), (0.2, 8192), (0.5, 4096), (1.0, 2048)):
        for _ in range(per_family):
            b = inputs.gaussian_polbeam(nb, 3 * fwhm, fwhm, float(rng.uniform(0, 0.1)), 0.0,
                                        nside_sky=ns)
            specs.append((t(np.stack(b.aBeams)), t(np.stack(b.bBeams)), nb,
                          float(rng.uniform(-np.pi, np.pi))))
            radii.append(b.get_rho_max())
    batch = DetectorBatch(specs)
    ctx = torch_ops._ctx(dev)
    stream = torch_ops._stream(dev)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    visits = 0
    for r in radii:
        check(lib.pisco_dev_query_disc(ctx, stream, ns, n, rat.data_ptr(), dect.data_ptr(), r,
                                       None, counts.data_ptr(), None))
        visits += int(counts.to(torch.int64).sum().item())
    g = torch.randn((len(specs), 2, n), device=dev)

    def tm(f):
        f()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_f = tm(lambda: batch.forward(It, Qt, Ut, rat, dect, pat, ns))
    t_a = tm(lambda: batch.adjoint_map(g, rat, dect, pat, ns))
    batch.close()
    det = 2 * len(specs) * n
    b_f = visits * 36 + len(specs) * n * 20
    b_a = visits * 48 + len(specs) * n * 20
    return {"workload": "256 detectors (128 pairs, FWHM 0.1/0.2/0.5/1.0 deg x 32 pairs, cap "
                        "3xFWHM, beam NSIDE 8192/8192/4096/2048), sky NSIDE 2048 power-law + 200 "
                        "point sources, %d samples CES" % n,
            "pair_visits": visits,
            "batched_forward": {"det_samples_per_s": det / t_f, "s": t_f,
                                "hbm_frac": b_f / t_f / HBM_PEAK},
            "batched_adjoint": {"det_samples_per_s": det / t_a, "s": t_a,
                                "hbm_frac": b_a / t_a / HBM_PEAK},
            "fwd_adj": {"det_samples_per_s": det / (t_f + t_a), "s": t_f + t_a,
                        "hbm_frac": (b_f + b_a) / (t_f + t_a) / HBM_PEAK}}


def map_maker_workload(dev, nside_sky, n=250_000, ndet=64, reps=3, with_cpu=True,
                       cpu_dets=8):
    """Binned map-making (SkyMap::accumulate_data + solve_map) of 64 detectors
    x 2.5e5 samples sharing one CES pointing, each detector with its own
    polarisation angle, into a sky of NSIDE 512.  Gather model: each
    detector-sample reads its pointing and datum (16 B); each touched pixel
    reads and writes its nine FP64 sums once per call (144 B); the solve
    reads every pixel's hit count and writes I, Q, U, hits (24 B per pixel)
    and reads the other eight sums of pixels with 3 or more hits (64 B).  The CPU
    oracle (sequential, as the reference's loops) runs on the first
    `cpu_dets` detectors."""
    import numpy as np
    import torch

    from pisco4cmb_amd import inputs
    from pisco4cmb_amd.mapmaker import DeviceSkyMap
    ra, dec, pa = inputs.ces_scan(n, t0=3600.0)
    rng = np.random.default_rng(65)
    pol = rng.uniform(-np.pi, np.pi, ndet).astype(np.float32)
    data = rng.normal(0, 30, (ndet, n)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rat, dect, pat, polt, datat = t(ra), t(dec), t(pa), t(pol), t(data)
    m = DeviceSkyMap(nside_sky, dev)

    def tm(f):
        f()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t_acc = tm(lambda: m.accumulate(rat, dect, pat, polt, datat, shared_pointing=True))
    t_sol = tm(m.solve)
    hits = m.solve()[3]
    touched = int((hits > 0).sum().item())
    solved = int((hits >= 3).sum().item())
    m.close()
    npix = 12 * nside_sky ** 2
    b_acc = ndet * n * 16 + touched * 144
    b_sol = npix * 24 + solved * 64
    res = {"workload": "map-making of %d detectors x %d samples (one CES pointing, per-detector "
                       "polarisation angles) into NSIDE %d" % (ndet, n, nside_sky),
           "touched_pixels": touched, "solved_pixels": solved,
           "accumulate": {"det_samples_per_s": ndet * n / t_acc, "s": t_acc,
                          "hbm_frac": b_acc / t_acc / HBM_PEAK},
           "solve": {"pixels_per_s": npix / t_sol, "s": t_sol,
                     "hbm_frac": b_sol / t_sol / HBM_PEAK}}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind
        o = oracle_bind.Oracle()
        c0 = time.time()
        AtA, AtD, _ = o.skymap_accumulate(nside_sky, ra, dec, pa, pol[:cpu_dets],
                                          data[:cpu_dets], shared_pointing=True)
        c1 = time.time()
        o.skymap_solve(nside_sky, AtA, AtD)
        c2 = time.time()
        res["cpu_oracle"] = {"accumulate_det_samples_per_s": cpu_dets * n / (c1 - c0),
                             "solve_pixels_per_s": npix / (c2 - c1), "cores": 1,
                             "sample": "first %d detectors, sequential as skymap.cpp"
                                       % cpu_dets}
    return res


def cpu_threads():
    """(threads, report): the CPUs this process may run on (its affinity
    mask, the host cores the job is given), next to nproc and the OpenMP
    limit of the environment, which the reference's OpenMP loop
    (convolver.cpp:38-41) would also obey."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:   # pragma: no cover - non-Linux
        allowed = nproc
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = allowed
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    quota = None
    try:   # cgroup v2 CPU quota, in CPUs
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    return threads, {"nproc": nproc, "affinity_cpus": allowed, "omp_num_threads": omp or None,
                     "cgroup_cpu_quota": quota}


def median_time(fn, reps=3):
    """One warm-up call, then the median of reps timed calls (seconds)."""
    fn(warm=True)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn(warm=False)
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2], ts


def cpu_baseline(args, I, Q, U, nside_sky, nside_beam, beam, ra, dec, pa, g, tod_gpu=None):
    """The CPU oracle (a C restatement of the reference Convolver,
    convolver.cpp:22-183, OpenMP over samples) on every CPU this process may
    use, one warm-up and the median of 3 runs:
      - value: forward + adjoint-map on the first cpu_samples samples of the
        headline workload (the metric's workload);
      - c1_full: configuration 1 in full (sky NSIDE 128, beam NSIDE 1024,
        cap 2.5 deg, 1e4 raster samples), forward;
      - c2_prefix: the first 1e5 samples of configuration 2 (sky NSIDE 256,
        beam NSIDE 2048, cap 1.5 deg, CES scan), forward.
    With the GPU TOD of the timed steps (tod_gpu = (a, b)), also the parity of
    that TOD on a prefix: relative RMS against the CPU Convolver and against
    the all-FP64 oracle, next to the CPU Convolver's own error against the
    FP64 oracle."""
    import numpy as np

    from pisco4cmb_amd import inputs
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind
    orc = oracle_bind.Oracle()
    threads, hosts = cpu_threads()
    m = min(args.cpu_samples, ra.size)
    beams = list(beam.aBeams) + list(beam.bBeams)

    def c3(warm):
        k = min(m, 2000) if warm else m
        orc.convolve(nside_sky, I, Q, U, nside_beam, beams, ra[:k], dec[:k], pa[:k],
                     nthreads=threads)
        orc.adjoint_map(nside_sky, nside_beam, beams, ra[:k], dec[:k], pa[:k], g[0, :k],
                        g[1, :k])
    t3, ts3 = median_time(c3)
    out = {"value": 2 * m / t3, "unit": "TOD samples/s", "cores": threads,
           "kind": "port",
           "sample": "first %d samples of the same workload, forward + adjoint-map, OpenMP "
                     "over samples on %d threads; one warm-up, median of 3 (%s s)"
                     % (m, threads, ", ".join("%.2f" % x for x in ts3)),
           "hosts": hosts}

    # configuration 1 in full: CPU plain, 1 detector pair's forward
    I1, Q1, U1, _ = inputs.iid_sky(128, seed=1)
    b1 = inputs.gaussian_polbeam(1024, 2.5, 1.0, 0.05, 0.01, nside_sky=128)
    r1, d1, p1 = inputs.raster_scan(10000)
    bl1 = list(b1.aBeams) + list(b1.bBeams)
    t1, ts1 = median_time(lambda warm: orc.convolve(128, I1, Q1, U1, 1024, bl1, r1, d1, p1,
                                                    nthreads=threads))
    out["c1_full"] = {"tod_samples_per_s": 2 * r1.size / t1, "s": t1,
                      "sample": "configuration 1 in full: %d raster samples, forward of the "
                                "pair" % r1.size, "runs_s": ts1}
    # configuration 2, first 1e5 samples
    I2, Q2, U2, _ = inputs.powerlaw_sky(256, seed=77, lmax=256)
    b2 = inputs.gaussian_polbeam(2048, 1.5, 0.5, 0.0, 0.0, nside_sky=256)
    r2, d2, p2 = inputs.ces_scan(100_000)
    bl2 = list(b2.aBeams) + list(b2.bBeams)
    t2, ts2 = median_time(lambda warm: orc.convolve(256, I2, Q2, U2, 2048, bl2, r2, d2, p2,
                                                    nthreads=threads))
    out["c2_prefix"] = {"tod_samples_per_s": 2 * r2.size / t2, "s": t2,
                        "sample": "first %d samples of configuration 2's CES scan, forward "
                                  "of the pair" % r2.size, "runs_s": ts2}
    if tod_gpu is not None:
        mp = min(m, 5000)
        ca, cb = orc.convolve(nside_sky, I, Q, U, nside_beam, beams, ra[:mp], dec[:mp], pa[:mp],
                              nthreads=threads)
        da, db = orc.convolve(nside_sky, I, Q, U, nside_beam, beams, ra[:mp], dec[:mp], pa[:mp],
                              precision64=True, nthreads=threads)

        def rr(x, y):
            x = np.concatenate([np.asarray(v, np.float64) for v in x])
            y = np.concatenate([np.asarray(v, np.float64) for v in y])
            return float(np.sqrt(np.mean((x - y) ** 2)) / np.sqrt(np.mean(y ** 2)))
        ga, gb = (np.asarray(v[:mp]) for v in tod_gpu)
        out["parity"] = {"samples": mp,
                         "gpu_vs_cpu_convolver_rel_rms": rr((ga, gb), (ca, cb)),
                         "gpu_vs_fp64_oracle_rel_rms": rr((ga, gb), (da, db)),
                         "cpu_convolver_vs_fp64_oracle_rel_rms": rr((ca, cb), (da, db)),
                         "tolerance": 1e-6}
    return out


if __name__ == "__main__":
    main()
