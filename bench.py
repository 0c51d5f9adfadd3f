"""H|psi> throughput on one MI355X: the 64-atom cell, FP64 vs mixed.

A step is one H|psi> application to all 256 bands of the 64-atom cell
(Si64 diamond 2x2x2 supercell, a = 20.52 bohr, ecut 40 Ry, npw 36,945,
90^3 FFT grid, 512 projectors, k = (0.1, 0, 0)), with psi already resident
in HBM.  Prints one JSON line (rank 0) with throughput, the roofline of the
dominant kernel (HIP events on the library's stream), the mixed-vs-FP64
energy error, FP64 parity against the CPU oracle and the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
RY_PER_HA = 2.0


TRAFFIC_PROFILE = "profiles/hpsi_si64_traffic.md"


def measured_traffic(kernel, prec, shapes=None):
    """HBM bytes per launch of a kernel class at the 64-atom workload, from the
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE scaled by what they read
    for a plain 1 GiB device copy); for the x-pencil kernel the variant with
    this run's z width and bands per workgroup.  None if the summary is
    absent."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        ent = json.loads(block)[kernel][prec]
        if shapes is not None:
            suffix = f", {shapes['pencil_z']}, {shapes['pencil_bands']}>"
            for name, t in ent.get("variants", {}).items():
                if name.endswith(suffix):
                    return t
        return ent["traffic_bytes"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


def kernel_model(name, prec, nb, npw, nplanes, plane, N, nch):
    """Algorithmic bytes (or flops) of one launch of a kernel class over nb
    bands: each declared input / output counted once at storage precision."""
    S = 16 if prec == "fp64" else 8
    R = 8 if prec == "fp64" else 4
    if name == "plane_g2r":        # psi gather + (ig, pos) int32 pairs, plane writes
        return "hbm", nb * (npw * (S + 8) + nplanes * plane * S)
    if name == "xpencil_vloc":     # plane read + write, V_loc once
        return "hbm", nb * 2 * nplanes * plane * S + N * R
    if name == "plane_r2g":        # plane read, psi read, hpsi write, indices, kinetic
        return "hbm", nb * (nplanes * plane * S + npw * (2 * S + 8)) + npw * 8
    if name == "proj_gemm":        # <p|psi>: 8 flop per complex MAC
        return "mfma", 8.0 * nch * npw * nb
    if name == "proj_apply":
        return "mfma", 8.0 * nch * npw * nb
    if name == "proj_reduce":
        return "hbm", None
    return "hbm", None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-bands", type=int, default=96,
                    help="bands in the 1-core CPU sample (about 12 s of CPU work at the 64-atom size)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--timing-steps", type=int, default=3)
    ap.add_argument("--serve-queries", type=int, default=1000,
                    help="effective-mass queries of the serving stream replayed at 50/s (0 = skip)")
    ap.add_argument("--serve-light", type=int, default=100,
                    help="queries of a second stream at --serve-light-rate, below the server's "
                         "capacity, for the unqueued latency (0 = skip)")
    ap.add_argument("--serve-light-rate", type=float, default=4.0)
    ap.add_argument("--serve-mid", type=int, default=200,
                    help="queries of a third stream at --serve-mid-rate, near the server's capacity "
                         "(0 = skip)")
    ap.add_argument("--serve-mid-rate", type=float, default=8.0)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--kp-mass", type=int, default=1,
                    help="1: the 8-atom cell's k.p masses in FP64 and mixed (kp_mass); 0 = skip")
    ap.add_argument("--only", choices=("fp64", "mixed"), default=None,
                    help="time one precision only (profiling runs)")
    ap.add_argument("--multik-steps", type=int, default=3,
                    help="timed steps of the 32-k-point batched H|psi> + dH/dk leg (0 = skip)")
    ap.add_argument("--facade", type=int, default=1,
                    help="time the C++ facade's one-band-per-call pattern (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="run a plain device copy of MB megabytes 5 times and exit "
                         "(calibrates rocprof FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU (--gpus 1); the H|psi> path shards over "
                                   "k-points and bands without collectives, and multi-GPU runs are not "
                                   "implemented"}))
        sys.exit(2)

    import torch
    from effmass_amd import synth
    from effmass_amd.hamiltonian import Hamiltonian

    t_setup = time.time()
    wl = synth.build("si64_cubic")
    cell, kp, psi_h = wl.cell, wl.kpts[0], wl.psi[0]
    nb, npw = psi_h.shape
    H = Hamiltonian(cell, device=0)
    H.load_kpoint(0, kp)
    dev = torch.device("cuda:0")
    psi64 = torch.from_numpy(psi_h).to(dev)
    psi32 = psi64.to(torch.complex64)
    h64 = torch.zeros_like(psi64)
    h32 = torch.zeros_like(psi32)
    op = H.at_k(0)
    setup_s = time.time() - t_setup

    def run(prec, steps):
        p, h = (psi64, h64) if prec == "fp64" else (psi32, h32)
        for _ in range(steps):
            op(p, h, prec=prec, overwrite=True)

    precs = (args.only,) if args.only else ("fp64", "mixed")
    # launch shapes picked by the autotuner on this cell, outside the timed region
    tuned = {}
    for prec in precs:
        H.autotune(0, nb, prec)
        tuned[prec] = H.shapes(prec)
    res = {}
    for prec in precs:
        run(prec, args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(prec, args.steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[prec] = {"ms_per_step": dt * 1e3 / args.steps, "bands_per_s": nb * args.steps / dt}

    if args.only:
        print(json.dumps({"only": args.only, **res[args.only]}), flush=True)
        return
    # energies: FP64 Rayleigh quotients of both paths
    op(psi64, h64, prec="fp64", overwrite=True)
    op(psi32, h32, prec="mixed", overwrite=True)
    E64 = H.rayleigh(psi64, h64).cpu().numpy()
    E32 = H.rayleigh(psi32, h32, prec="mixed").cpu().numpy()
    dE_ry = float(np.abs(E32 - E64).max() * RY_PER_HA)

    # per-kernel durations with HIP events on the library's stream
    geo = dict(nb=nb, npw=npw, nplanes=len(np.unique(kp.gidx[:, 0] % cell.n[0])),
               plane=cell.n[1] * cell.n[2], N=int(np.prod(cell.n)), nch=cell.nchannels())
    kernels = []
    for prec in ("fp64", "mixed"):
        H.reset_timing()
        H.set_timing(True)
        run(prec, args.timing_steps)
        torch.cuda.synchronize()
        for name in ("plane_g2r", "xpencil_vloc", "plane_r2g", "proj_gemm", "proj_reduce", "proj_apply"):
            ms, cnt = H.timing(name)
            if cnt == 0:
                continue
            avg_s = ms / cnt * 1e-3
            kind = "fp64" if prec == "fp64" else "fp32"
            # bands per launch: the FFT passes run in band chunks when the
            # workspace is capped, the projector kernels on all bands
            g = dict(geo, nb=nb * args.timing_steps // cnt)
            bound, amount = kernel_model(name, "fp64" if prec == "fp64" else "fp32", **g)
            ent = {"kernel": name, "prec": prec, "avg_us": avg_s * 1e6, "launches": cnt,
                   "share_of_step": ms / cnt / res[prec]["ms_per_step"], "bound": bound}
            if amount is not None:
                if bound == "hbm":
                    ach = amount / avg_s / 1e9
                    ent.update(achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s", frac=ach / HBM_PEAK_GBS,
                               algorithmic_bytes=amount)
                else:
                    ach = amount / avg_s / 1e12
                    pk = MFMA_PEAK_TFLOPS[kind]
                    ent.update(achieved=ach, peak=pk, unit="TFLOP/s", frac=ach / pk, algorithmic_flops=amount)
            kernels.append(ent)
        H.set_timing(False)
    dom = max((k for k in kernels if k["prec"] == "mixed" and "frac" in k), key=lambda k: k["share_of_step"],
              default=None)
    roofline = None                       # needs the per-kernel timing steps (--timing-steps > 0)
    if dom is not None:
        traffic = measured_traffic(dom["kernel"], "mixed",
                                   tuned.get("mixed") if dom["kernel"] == "xpencil_vloc" else None)
        roofline = {"bound": dom["bound"], "achieved": dom["achieved"], "peak": dom["peak"], "unit": dom["unit"],
                    "frac": dom["frac"], "traffic": traffic, "kernel": dom["kernel"], "prec": "mixed",
                    "traffic_source": TRAFFIC_PROFILE if traffic is not None else None}
        if dom["bound"] == "hbm":
            # what a plain device copy sustains on this device in this run
            # (bytes read + written per second), and the dominant kernel's
            # measured traffic against it
            copy_gbs = device_copy_gbs()
            roofline["device_copy_gbs"] = copy_gbs
            if traffic is not None:
                rate = traffic / (dom["avg_us"] * 1e-6) / 1e9
                roofline["traffic_gbs"] = rate
                roofline["traffic_frac_of_copy"] = rate / copy_gbs

    # CPU baseline: the FP64 oracle (reference loop structure) on host cores
    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        nsb = min(args.cpu_bands, nb)
        sample = psi_h[:nsb]
        oracle.hpsi(cell, kp, sample[:1])                      # warm-up
        t0 = time.perf_counter()
        ref = oracle.hpsi(cell, kp, sample)
        t1 = time.perf_counter() - t0
        nthr = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        oracle.hpsi(cell, kp, psi_h[:4 * nthr], nthreads=nthr)
        tn = time.perf_counter() - t0
        split = _cpu_split(oracle, cell, kp, psi_h[:min(32, nb)], t1 / nsb)
        # value: the reference's algorithm on one core with the reference's
        # FFT library (pocketfft via numpy) for the local part and the
        # oracle's loop-structured nonlocal part; the pure-oracle rate (its
        # own slower FFT) beside it
        cpu = {"value": split.pop("value_with_pocketfft_local"), "unit": "bands/s", "cores": 1, "kind": "port",
               "sample": (f"64-atom cell, FP64, 1 thread: local part (scatter, 3x1D pocketfft, V, 3x1D "
                          f"pocketfft, gather) and the oracle's nonlocal part on {min(32, nb)} bands; "
                          f"the whole oracle on {nsb} bands took {t1:.1f} s"),
               "value_oracle_only": nsb / t1,
               "value_all_cores_oracle": 4 * nthr / tn, "cores_all": nthr,
               "cpu_model": _cpu_model(), **split}
        op(psi64, h64, prec="fp64", overwrite=True)
        got = h64[:nsb].cpu().numpy()
        rel = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
        op(psi32, h32, prec="mixed", overwrite=True)
        got32 = h32[:nsb].cpu().numpy()
        rel32 = np.linalg.norm(got32 - ref, axis=1) / np.linalg.norm(ref, axis=1)
        parity = {"fp64_max_rel_err_vs_oracle": float(rel.max()),
                  "mixed_max_rel_err_vs_oracle": float(rel32.max()), "bands_checked": nsb}

    H.close()
    # the C++ facade driven as the reference's callers drive it (one band
    # per call on host spans) beside its batched and device forms
    facade = facade_leg([("si2_fcc", synth.build("si2_fcc")), ("si64_cubic", wl)]) if args.facade else None
    # 32 k-points of the 64-atom cell in batched launches, graph replays
    multik = multik_leg(args.multik_steps) if args.multik_steps > 0 else None
    # k.p effective masses of the 8-atom cell, mixed vs FP64
    kpm = kp_mass_leg() if args.kp_mass else None
    # the vendor library on GEMMs of the same shapes, for reference
    libgemm = library_gemm_leg(npw, nb, cell.nchannels(), {k["kernel"] + "/" + k["prec"]: k["avg_us"] for k in kernels})

    # serving: Poisson(50/s) effective-mass queries over 2-64 atom cells
    serving_stats = None
    if args.serve_queries > 0 or args.serve_light > 0 or args.serve_mid > 0:
        from effmass_amd import serving
        torch.cuda.empty_cache()
        serving_stats = {"workload": ("Poisson arrivals; cells of 2/8/16/32/64 atoms at 135 bohr^3/atom, "
                                      "ecut 30/40/50 Ry, 16-256 bands, 1-4 k-points; m* of band nb/2 at "
                                      "every k from the k.p sum over the nb lowest states, all converged "
                                      "(Chebyshev-filtered subspace iteration from a plane-wave Rayleigh-Ritz "
                                      "start, max(16, nb/4) buffer rows, mixed precision); latency = "
                                      "arrival to answer"),
                         "slots": args.serve_slots}
        for key, n, rate, seed in (("rate_50", args.serve_queries, 50.0, 2024),
                                   ("below_capacity", args.serve_light, args.serve_light_rate, 2025),
                                   ("near_capacity", args.serve_mid, args.serve_mid_rate, 2026)):
            if n <= 0:
                continue
            qs = serving.make_stream(n, rate=rate, seed=seed)
            _, st = serving.serve(qs, slots=args.serve_slots, prec="mixed", timeout=1200,
                                  progress=_progress(f"serving {key}"))
            st["rate_per_s"] = rate
            serving_stats[key] = st
        # cost model of the stream: each size class alone (one k-point, 128
        # bands, 40 Ry), and the rate the device's kernel time allows for the
        # stream's mix (uniform classes and cutoffs, mean nb 136, mean 2.5
        # k-points per query), were the kernels of concurrent queries packed
        # back to back
        br = serving.class_breakdown(prec="mixed", ecut_ry=40.0, nb=128, reps=2)
        per_part_ms = float(np.mean([v["kernel_ms"] for v in br.values()])) * (136.0 / 128.0)
        serving_stats["class_breakdown"] = {"workload": "one query alone per size class: 1 k-point, 128 bands, "
                                                        "40 Ry, mixed; times in s, kernel time in ms",
                                            "classes": br,
                                            "stream_kernel_ms_per_query_estimate": per_part_ms * 2.5,
                                            "device_bound_queries_per_s_estimate": 1e3 / (per_part_ms * 2.5)}

    out = {
        "metric": "Hψ bands/s, 64-atom cell, mixed vs FP64 at 1e-6 Ry; p50/p99 m* query latency",
        "value": res["mixed"]["bands_per_s"],
        "unit": "bands/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["mixed"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "precision": "f32 FFT / GEMM (complex64 psi), f64 D contraction, kinetic term, Rayleigh quotients",
        "data": "synthetic (seeded SplitMix64, effmass_amd/synth.py workload si64_cubic)",
        "config": {"workload": "Si64 2x2x2 diamond supercell, a=20.52 bohr, ecut 40 Ry, k=(0.1,0,0), "
                               "npw 36945, FFT 90^3, 256 bands, 512 projectors; step = H|psi> on all bands",
                   "bands": nb, "npw": npw, "grid": list(cell.n), "projectors": cell.nchannels()},
        "fp64_bands_per_s": res["fp64"]["bands_per_s"],
        "fp64_ms_per_step": res["fp64"]["ms_per_step"],
        "mixed_over_fp64": res["mixed"]["bands_per_s"] / res["fp64"]["bands_per_s"],
        "energy_err_mixed_vs_fp64_ry": dE_ry,
        "parity": parity,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "tuned_launch_shapes": tuned,
        "multi_k": multik,
        "kp_mass": kpm,
        "library_gemm": libgemm,
        "facade": facade,
        "m_star_latency": _latency_summary(serving_stats),
        "serving": serving_stats,
        "setup_s": setup_s,
    }
    print(json.dumps(out), flush=True)


def device_copy_gbs(mib=1024, reps=5):
    """Bytes read + written per second by torch's device-to-device copy of
    `mib` MiB (HIP events, best of `reps`): the rate a streaming kernel can
    reach on this device."""
    import torch
    a = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del a, b
    torch.cuda.empty_cache()
    return 2 * mib * (1 << 20) / (best * 1e-3) / 1e9


def library_gemm_leg(npw, nb, nch, ours, reps=5):
    """torch.matmul (rocBLAS cgemm / zgemm) on random operands of the shapes
    of the two projector GEMMs of the step, beside this library's kernels:
    the apply out[b][g] += sum_c C[c][b] P[c][g] (M = npw, N = nb, K = nch)
    and the projection <p|psi> = conj(P) psi^T (M = nch, N = nb, K = npw).
    Algorithmic flops 8 M N K as for the library's kernels."""
    import torch
    dev = "cuda:0"
    out = {"workload": f"random operands, apply {npw} x {nb} x {nch}, projection {nch} x {nb} x {npw}",
           "library": "torch.matmul / addmm (rocBLAS)"}
    flops = 8.0 * npw * nb * nch
    g = torch.Generator(device=dev).manual_seed(5)
    for prec, dt, peak in (("fp64", torch.complex128, 78.6), ("mixed", torch.complex64, 157.3)):
        P = torch.randn(nch, npw, dtype=dt, device=dev, generator=g)
        Ct = torch.randn(nb, nch, dtype=dt, device=dev, generator=g)          # C^T, contiguous
        psiT = torch.randn(npw, nb, dtype=dt, device=dev, generator=g)        # psi^T, contiguous
        h = torch.zeros(nb, npw, dtype=dt, device=dev)
        Pc = P.conj().resolve_conj()

        def timed(fn):
            for _ in range(2):
                fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) * 1e3 / reps
        t_apply = timed(lambda: h.addmm_(Ct, P))
        t_proj = timed(lambda: torch.matmul(Pc, psiT))
        ent = {}
        for name, t in (("apply", t_apply), ("projection", t_proj)):
            key = "proj_apply" if name == "apply" else "proj_gemm"
            mine = ours.get(f"{key}/{prec}")
            ent[name] = {"library_us": t, "library_frac": flops / t / 1e6 / peak, "ours_us": mine,
                         "speedup": (t / mine) if mine else None}
        out[prec] = ent
        del P, Ct, psiT, h, Pc
    torch.cuda.empty_cache()
    return out


def multik_leg(steps, warmup=1):
    """H|psi> and dH/dk_{x,y,z}|psi> for 32 k-points (4x4x2 grid, ragged npw
    36,816-37,073) x 128 bands of the 64-atom cell: effm_hpsi_dhdk_multi
    (one projection GEMM for both, up to 16 k-points per launch), replayed
    as a captured hipGraph; psi is seeded device noise (throughput does not
    depend on the values).  Also times the same work as effm_hpsi_multi +
    effm_dhdk_multi (graph and eager) and as one call per k-point."""
    import torch
    from effmass_amd import synth
    from effmass_amd.hamiltonian import GradHPsi, Hamiltonian
    wl = synth.build("si64_kgrid", with_psi=False)
    H = Hamiltonian(wl.cell, device=0)
    for ik, kp in enumerate(wl.kpts):
        H.load_kpoint(ik, kp)
    ks = list(range(len(wl.kpts)))
    nb = wl.nb
    gen = torch.Generator(device="cuda:0").manual_seed(2024)
    res = {"workload": "Si64 cell, 4x4x2 k-grid (32 k, npw %d-%d), 128 bands per k; step = H|psi> + "
                       "dH/dk_x,y,z|psi> on every k" % (min(k.npw for k in wl.kpts), max(k.npw for k in wl.kpts)),
           "kpoints": len(ks), "bands_per_k": nb}
    for prec, dt in (("mixed", torch.complex64), ("fp64", torch.complex128)):
        psis = [torch.randn(nb, kp.npw, dtype=dt, device="cuda:0", generator=gen) for kp in wl.kpts]
        hs = [torch.empty_like(p) for p in psis]
        gr = GradHPsi([torch.empty_like(p) for p in psis], [torch.empty_like(p) for p in psis],
                      [torch.empty_like(p) for p in psis])

        def step(graph=True):
            H.apply_with_gradient_multi(ks, psis, hs, gr, prec=prec, overwrite=True, graph=graph)

        def separate(graph=True):
            H.apply_multi(ks, psis, hs, prec=prec, overwrite=True, graph=graph)
            H.gradient_multi(ks, psis, gr, prec=prec, overwrite=True, graph=graph)

        def per_k():
            for ik in ks:
                H.at_k(ik)(psis[ik], hs[ik], prec=prec, overwrite=True)
                H.gradient().at_k(ik)(psis[ik], GradHPsi(gr.x[ik], gr.y[ik], gr.z[ik]), prec=prec,
                                      overwrite=True)

        def timed(fn, n):
            for _ in range(warmup):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / n

        t_graph = timed(step, steps)
        t_eager = timed(lambda: step(False), steps)
        t_sep = timed(separate, steps)
        t_loop = timed(per_k, steps)
        # parity of the timed path: a graph replay of the batch against
        # single-k effm_hpsi + effm_dhdk calls on the first bands of two
        # k-points (one in each launch group of 16)
        step(True)
        nchk = 4
        worst = 0.0
        for ik in (0, len(ks) - 1):
            p = psis[ik][:nchk].contiguous()
            sh = torch.zeros_like(p)
            sg = [torch.zeros_like(p) for _ in range(3)]
            H.at_k(ik)(p, sh, prec=prec, overwrite=True)
            H.gradient().at_k(ik)(p, GradHPsi(*sg), prec=prec, overwrite=True)
            for got, ref in [(hs[ik], sh), (gr.x[ik], sg[0]), (gr.y[ik], sg[1]), (gr.z[ik], sg[2])]:
                d = (got[:nchk] - ref).abs().pow(2).sum(1).sqrt() / ref.abs().pow(2).sum(1).sqrt()
                worst = max(worst, float(d.max()))
        parity = {"max_rel_err_batch_graph_vs_single_k": worst, "kpoints_checked": [0, len(ks) - 1],
                  "bands_checked": nchk, "outputs": "H|psi>, dH/dk_x,y,z|psi>"}
        # per-kernel split of one eager batched step (HIP events)
        H.reset_timing()
        H.set_timing(True)
        step(False)
        torch.cuda.synchronize()
        split = {}
        for name in ("plane_g2r", "xpencil_vloc", "plane_r2g", "proj_gemm", "proj_reduce", "proj_apply"):
            ms, cnt = H.timing(name)
            if cnt:
                split[name] = {"ms": ms, "launches": cnt}
        H.set_timing(False)
        applies = len(ks) * nb
        res[prec] = {"ms_per_step": t_graph * 1e3, "k_band_steps_per_s": applies / t_graph,
                     "eager_batched_ms": t_eager * 1e3, "separate_calls_graph_ms": t_sep * 1e3,
                     "per_k_calls_ms": t_loop * 1e3, "kernel_ms": split, "parity": parity}
        del psis, hs, gr
        torch.cuda.empty_cache()
    H.close()
    return res


def kp_mass_leg():
    """The 8-atom cubic cell (a = 10.26 bohr, ecut 50 Ry, 48^3, npw 6451) at
    Gamma: the 200 lowest states by the Chebyshev-filtered subspace iteration
    (plane-wave start, 48 buffer rows; residual 1e-8 Ha in FP64, 2e-5 mixed),
    then 1/m* of bands 4 and 5 from the k.p sum over the 200 states (FP64
    sums and denominators in both).  Reported: the tensors, the error of the
    mixed path per tensor element against the FP64 path (whose masses equal
    the oracle's on the same eigenpairs to 1e-9, tests/test_gpu_effmass.py),
    and the device seconds of each part."""
    import torch
    from effmass_amd import effmass, eigensolver, synth
    from effmass_amd.hamiltonian import Hamiltonian

    wl = synth.build("si8_cubic", nb=2)
    cell, kp = wl.cell, wl.kpts[0]
    bands, nb, rows = [4, 5], 200, 248
    out = {"workload": "si8_cubic: 8-atom cubic Si, a=10.26 bohr, ecut 50 Ry, npw %d, FFT 48^3, Gamma; "
                       "200 states, 1/m* of bands 4, 5 by the k.p sum over them" % kp.npw}
    res = {}
    for prec in ("fp64", "mixed"):
        H = Hamiltonian(cell)
        H.load_kpoint(0, kp)
        kin = torch.from_numpy(kp.kin).cuda()
        op = H.at_k(0)
        wdt = torch.complex128 if prec == "fp64" else torch.complex64

        def apply(V, HV):
            op(V, HV, prec=prec, overwrite=True)
        # untimed warm-up of the same calls on a small block (library
        # kernels, solver and allocator first calls)
        Xw = eigensolver.plane_wave_start(apply, kin, 24, work_dtype=wdt, eigh=eigensolver.eigh_host)
        eigensolver.chfsi(apply, Xw, kin, nconv=8, tol=1e-3, maxiter=2, work_dtype=wdt, gram=H.overlap,
                          degree=8, eigh=eigensolver.eigh_host, lincomb=H.lincomb, buffer_degree="wanted")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        X0 = eigensolver.plane_wave_start(apply, kin, rows, work_dtype=wdt, eigh=eigensolver.eigh_host)
        r = eigensolver.chfsi(apply, X0, kin, nconv=nb, tol=1e-8 if prec == "fp64" else 2e-5, maxiter=300,
                              work_dtype=wdt, gram=H.overlap, degree=30, eigh=eigensolver.eigh_host,
                              lincomb=H.lincomb, buffer_degree="wanted")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        m = effmass.effective_mass(H, 0, kp, cell.A, r.X[:nb].contiguous(), r.E[:nb], bands, prec=prec)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        H.close()
        res[prec] = m
        out[prec] = {"solve_s": t1 - t0, "kp_s": t2 - t1, "iters": r.iters, "converged": r.converged,
                     "inv_mass": m.inv_mass.tolist(), "masses": [list(map(float, v)) for v in m.masses()]}
    d = np.abs(res["mixed"].inv_mass - res["fp64"].inv_mass)
    out["mixed_vs_fp64_abs_err_inv_mass"] = d.tolist()
    out["mixed_vs_fp64_max_rel_err_inv_mass"] = float(d.max() / np.abs(res["fp64"].inv_mass).max())
    mr = [np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b))
          for a, b in zip(res["mixed"].masses(), res["fp64"].masses())]
    out["mixed_vs_fp64_max_rel_err_m_star"] = float(max(x.max() for x in mr))
    return out


def _cpu_split(oracle, cell, kp, psi, per_band_s):
    """Local / nonlocal split of the oracle's per-band time, and the local
    part redone with numpy.fft (pocketfft, the FFT library the reference
    calls) in the reference's order: scatter, backward 1D FFTs along k, j,
    then i (unnormalised), V_loc multiply, forward 1D FFTs with 1/n each,
    gather (src/math/fft.cppm:60-85, hamiltonian_callable.cpp:138-182).
    The oracle's own FFT is a plain recursive mixed-radix FFT and slower
    than pocketfft, so the pocketfft figure is the closer estimate of the
    reference's local cost on this host."""
    nb = psi.shape[0]
    t0 = time.perf_counter()
    oracle.hpsi(cell, kp, psi, nonlocal_only=True)
    nl = (time.perf_counter() - t0) / nb
    n1, n2, n3 = cell.n
    V = cell.vloc.reshape(cell.n)
    g = kp.gidx
    idx = ((g[:, 0] % n1) * n2 + (g[:, 1] % n2)) * n3 + (g[:, 2] % n3)
    out = np.zeros_like(psi)
    t0 = time.perf_counter()
    for b in range(nb):
        grid = np.zeros(n1 * n2 * n3, np.complex128)
        grid[idx] = psi[b]
        grid = grid.reshape(cell.n)
        for ax in (2, 1, 0):
            grid = np.fft.ifft(grid, axis=ax, norm="forward")       # e^{+i}, no 1/n
        grid *= V
        for ax in (2, 1, 0):
            grid = np.fft.fft(grid, axis=ax, norm="forward")        # e^{-i}, 1/n
        out[b] = kp.kin * psi[b] + grid.reshape(-1)[idx]
    loc_np = (time.perf_counter() - t0) / nb
    loc_orc = max(per_band_s - nl, 0.0)
    return {"split_per_band_ms": {"oracle_local": loc_orc * 1e3, "oracle_nonlocal": nl * 1e3,
                                  "pocketfft_local": loc_np * 1e3},
            "value_with_pocketfft_local": 1.0 / (loc_np + nl),
            "split_sample": f"{nb} bands, 1 thread; numpy {np.__version__} pocketfft for the local part"}


def write_dump(path, cell, kp, psi):
    """The binary cell / k-point / band dump effmass_amd/host/facade_bench.cpp reads."""
    import struct
    i32 = lambda *v: struct.pack("<%di" % len(v), *v)
    f64 = lambda a: np.ascontiguousarray(a, dtype="<f8").tobytes()
    with open(path, "wb") as f:
        f.write(b"EFFMDUMP" + i32(1) + i32(*cell.n) + f64(np.asarray(cell.A).ravel()) + f64(cell.vloc))
        f.write(i32(cell.natom) + np.ascontiguousarray(cell.species, "<i4").tobytes() + f64(cell.tau))
        f.write(i32(len(cell.species_list)))
        for sp in cell.species_list:
            f.write(i32(sp.lmax, sp.nbeta) + np.ascontiguousarray(sp.beta_l, "<i4").tobytes() + f64(sp.lam))
            f.write(i32(sp.nq) + struct.pack("<2d", sp.dq, sp.qmax) + f64(sp.beta_tab))
            f.write(i32(1) + f64(sp.dbeta_tab) if sp.dbeta_tab is not None else i32(0))
        f.write(i32(kp.npw) + f64(kp.K) + f64(kp.kin) + np.ascontiguousarray(kp.gidx, "<i4").tobytes()
                + f64(kp.kfrac))
        f.write(i32(psi.shape[0]) + np.ascontiguousarray(psi, "<c16").tobytes())


def facade_leg(workloads):
    """ms per band of the C++ facade (effm_hamiltonian.hpp) on host spans,
    one band per call (the reference's pattern, test_hpsi_eigen.cpp:88) and
    batched, and on device pointers, for each workload."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "effmass_amd", "host", "_build", "facade_bench")
    if not os.path.exists(exe):
        return {"error": "facade_bench not built (__graft_entry__.build())"}
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, wl in workloads:
            path = os.path.join(d, name + ".dump")
            write_dump(path, wl.cell, wl.kpts[0], wl.psi[0])
            r = subprocess.run([exe, path, "3"], capture_output=True, text=True, timeout=600)
            out[name] = json.loads(r.stdout) if r.returncode == 0 else {"error": r.stderr.strip()[-300:]}
    return out


def _latency_summary(st):
    """p50/p99 of the 50/s stream (and of the below- and near-capacity streams)."""
    if not st:
        return None
    out = {}
    for key in ("rate_50", "below_capacity", "near_capacity"):
        if key in st:
            out[key] = {k: st[key][k] for k in ("p50_s", "p99_s", "queries_per_s", "queries", "rate_per_s")}
    return out


def _progress(tag):
    """A progress line on stderr about every 20 s (long serving streams)."""
    state = {"t": time.perf_counter()}

    def f(done, total):
        now = time.perf_counter()
        if now - state["t"] >= 20.0:
            state["t"] = now
            print(f"{tag}: {done}/{total} answered", file=sys.stderr, flush=True)
    return f


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
