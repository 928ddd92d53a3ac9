"""Benchmark: MB-pol 2B+3B energy+gradient of 2048 FP64 waters on one MI355X.

A step is one full Energy(true)-equivalent 2B+3B pass over the periodic 2048-water box
(edge 39.4 A, cutoffs 9 / 7 A): FixMonomerCoordinates, the dimer and trimer lists (cell-list
builder), the polynomial evaluation of every evaluated n-mer, gradient scatter and energy /
virial reduction.  Coordinates are resident in HBM before timing starts (a device-to-device copy
into the engine's working buffer is part of each step); results stay on the device.

Throughput is listed dimers + trimers per second (the n-mers the reference's Get2B / Get3B walk).

The JSON line also carries
  roofline:     the polynomial launch (the dominant kernel: the MB-pol 3B and 2B polynomials in
                one launch with the 3B chain rule in its 3B blocks, k_poly_eg_pair_chain3; without
                the chain rule when fuse_chain is off; the 3B one alone when the pairing is off)
                against MI355X FP64 vector peak, with the reference's algorithmic FLOPs per
                evaluated trimer (81,549, plus 1,094 for the chain rule) and dimer (27,208),
                counted from the reference's code (mbx_amd/tables/flops.json), timed with HIP events
                on the engine stream in a stage-timing pass of the same workload; traffic from the
                committed rocprofv3 PMC summary when present.  config.polynomial_alone times the
                launch without the chain rule beside it;
  cpu_baseline: the reference's own compiled 2B/3B leaves (oracle/_ref, OpenMP driver) on this
                host's cores, on the same workload, in two dispatch modes: the reference's own
                batch of 8 n-mers per polynomial call and a best-effort batch of 1024; the faster
                one is the baseline, both are reported with the CPU model.
Also in "config": the same step with the gradient copied to the host every step
(with_grad_d2h; PCIe-inclusive, so never the headline value) and the drop-in path a C / Fortran /
Python caller takes (drop_in: get_energy_g_ = SetRealXyz -> Energy(true) -> GetRealGrads of the
C++ System, 1B + 2B + 3B, host arrays in and out).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "2B+3B dimers+trimers/sec (E+F, FP64) at 2048 H2O; batched vs plain speedup"
FP64_VECTOR_PEAK = 78.6e12  # MI355X FP64 vector, FLOP/s
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_poly3b_water2048.txt")
LIST_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_lists_water2048.txt")
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s


def host_cores():
    """Threads for the CPU baseline: the cores this process may run on, capped by OMP_NUM_THREADS
    where the environment sets it (the GPU box gives one GPU's process a 16-thread share and sets
    OMP_NUM_THREADS=16; without either, 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(n, cap))


def host_cpu():
    """(model name, physical cores, logical CPUs) of this host from /proc/cpuinfo."""
    model, cores, logical = "unknown", set(), 0
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                logical += 1
            elif k == "model name":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
                cores.add((phys, core))
    except OSError:
        pass
    return model, len(cores) or logical, logical


def read_traffic(path=TRAFFIC_FILE, key="traffic_bytes_per_launch:"):
    try:
        for line in open(path):
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(xyz, box, c2, c3, seconds):
    """Reference leaves on the host (kind 'reference'), in the reference's batch-of-8 dispatch and
    a best-effort batch of 1024, each timed for about `seconds`; the faster mode is the baseline.
    The oracle port when the reference leaves are absent."""
    import oracle

    cores = host_cores()
    model, phys, logical = host_cpu()
    modes = {}
    r = None
    for batch in (8, 1024):
        reps, t_total = 0, 0.0
        if oracle.ref_available():
            kind = "reference"
            run = lambda: oracle.ref_water_2b3b(xyz, box, c2, c3, batch=batch, nthreads=cores)
        else:
            kind = "port"
            os.environ["OMP_NUM_THREADS"] = str(cores)
            run = lambda: oracle.water_system(xyz, box, c2, c3)
        r = run()  # warm-up
        while t_total < seconds or reps < 2:
            t0 = time.perf_counter()
            r = run()
            t_total += time.perf_counter() - t0
            reps += 1
        n = r["n2"] + r["n3"]
        modes["batch%d" % batch] = {"value": n * reps / t_total, "calls": reps, "cpu_s": t_total,
                                    "ms_per_call": t_total / reps * 1e3}
        if kind == "port":
            break
    best = max(modes, key=lambda m: modes[m]["value"])
    return dict(value=modes[best]["value"], unit="dimers+trimers/s", cores=cores, kind=kind,
                cpu_model=model, host_physical_cores=phys, host_logical_cpus=logical, mode=best, modes=modes,
                sample="%d full 2B+3B E+F calls on the 2048-water box (%.1f s of CPU time, %d threads; "
                       "the faster of the reference's batch-of-8 dispatch and a batch of 1024: %s)"
                       % (modes[best]["calls"], modes[best]["cpu_s"], cores, best)), r


def drop_in_leg(xyz, box, c2, c3, steps, tmpdir):
    """The reference-compatible entry points of a C / Fortran / Python caller: initialize_system_py_
    once, then get_energy_g_ per step (SetRealXyz -> Energy(true) -> GetRealGrads: 1B + 2B + 3B,
    host arrays in and out, one synchronous call)."""
    import ctypes

    import numpy as np

    from mbx_amd import engine

    L = engine.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    cpp = ctypes.POINTER(ctypes.c_char_p)
    L.initialize_system_py_.argtypes = [dp, ip, cpp, cpp, ip, ctypes.c_char_p]
    L.get_energy_g_.argtypes = [dp, ip, dp, dp]
    n = len(xyz) // 9
    js = os.path.join(tmpdir, "bench_mbx.json")
    with open(js, "w") as f:
        json.dump({"MBX": {"box": [float(v) for v in box], "twobody_cutoff": c2, "threebody_cutoff": c3}}, f)
    c = np.ascontiguousarray(xyz)
    L.initialize_system_py_(c.ctypes.data_as(dp), (ctypes.c_int * n)(*([3] * n)),
                            (ctypes.c_char_p * (3 * n))(*([b"O", b"H", b"H"] * n)),
                            (ctypes.c_char_p * n)(*([b"h2o"] * n)), ctypes.byref(ctypes.c_int(n)), js.encode())
    e = ctypes.c_double()
    g = np.zeros(len(xyz))
    nat = ctypes.c_int(3 * n)
    for _ in range(3):  # sizing, eager, then graph replays (the System runs its engine in graph mode)
        L.get_energy_g_(c.ctypes.data_as(dp), ctypes.byref(nat), ctypes.byref(e), g.ctypes.data_as(dp))
    t0 = time.perf_counter()
    for _ in range(steps):
        L.get_energy_g_(c.ctypes.data_as(dp), ctypes.byref(nat), ctypes.byref(e), g.ctypes.data_as(dp))
    dt = (time.perf_counter() - t0) / steps
    L.finalize_system_()
    return {"ms_per_call": dt * 1e3, "energy_1b_2b_3b": e.value,
            "call": "get_energy_g_: SetRealXyz -> Energy(true) (1B + 2B + 3B) -> GetRealGrads, host arrays, "
                    "synchronous"}, dt, g


def batched_vs_plain(parts, kinds, box, c2, c3, steps, warmup):
    """One batched compute over every configuration (one context, mbxg_set_batch) against one
    plain compute per configuration (one warmed-up context each), coordinates resident, HIP
    graph mode on both sides, no host wait until the end."""
    import numpy as np
    import torch

    from mbx_amd import engine

    sizes = [len(k) for k in kinds]
    off = np.concatenate([[0], np.cumsum(sizes)])
    big = engine.Engine()
    big.set_topology(np.concatenate(kinds))
    big.set_box(box)
    big.set_cutoffs(c2, c3)
    big.set_batch(off)
    big.set_coords(np.concatenate(parts))
    small = []
    for x, k in zip(parts, kinds):
        e = engine.Engine()
        e.set_topology(k)
        e.set_box(box)
        e.set_cutoffs(c2, c3)
        e.set_coords(x)
        small.append(e)
    big.compute()
    schedule = big.autotune(reps=2)
    big.set_graph(True)
    for e in small:
        e.set_graph(True)

    def timed(fn, sync):
        for _ in range(warmup):
            fn()
        sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        sync()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_b = timed(big.compute, big.synchronize)

    def run_plain():
        for e in small:
            e.compute()

    t_p = timed(run_plain, lambda: [e.synchronize() for e in small])
    c = big.counts()
    n_listed = int(c[0] + c[2])
    assert n_listed == sum(int(e.counts()[0] + e.counts()[2]) for e in small)
    e2, e3, _ = big.batch_results()
    dev = max(abs(e2[k] - small[k].results(grad=False)[0]) / max(abs(e2[k]), 1e-300) for k in range(len(small)))
    for e in small:
        e.close()
    big.close()
    return {"configs": len(parts), "molecules": int(off[-1]), "cutoff2b_A": c2, "cutoff3b_A": c3,
            "listed_nmers": n_listed, "evaluated_dimers": int(c[1]), "evaluated_trimers": int(c[3]),
            "ms_batched": t_b * 1e3, "ms_plain": t_p * 1e3, "batched_nmers_per_s": n_listed / t_b,
            "plain_nmers_per_s": n_listed / t_p, "speedup": t_p / t_b, "max_rel_dev_e2b_vs_plain": dev,
            "schedule": schedule,
            "plain": "one warmed-up context per configuration (HIP graph mode), one compute each, no host wait "
                     "until the end",
            "batched": "one context over all configurations (HIP graph mode), one compute"}


def batched_leg(steps, warmup):
    """512 gas-phase water clusters (16..512 waters), cutoffs 9 / 7 A."""
    import numpy as np

    from mbx_amd import configs

    cfg = configs.C3
    parts, sizes = configs.cluster_batch(cfg["nconf"], cfg["nmin"], cfg["nmax"], cfg["seed"])
    out = {"workload": cfg["name"]}
    out.update(batched_vs_plain(parts, [np.zeros(n, int) for n in sizes], None, cfg["cutoff2b"], cfg["cutoff3b"],
                                steps, warmup))
    return out


def pimd_leg(steps, warmup):
    """32 path-integral beads of a periodic 57 CH4 + 513 H2O box (all five n-mer families)."""
    from mbx_amd import configs

    cfg = configs.C5
    xyz, kinds, box = configs.mixed_box(cfg["n_ch4"], cfg["n_h2o"], seed=cfg["seed"])
    parts = configs.beads(xyz, cfg["nbeads"])
    out = {"workload": cfg["name"]}
    out.update(batched_vs_plain(parts, [kinds] * len(parts), box, cfg["cutoff2b"], cfg["cutoff3b"], steps, warmup))
    return out


def differentiation_leg(steps, warmup):
    """Energy + force loss over 4096 water dimers and 4096 water trimers, differentiated with
    respect to coordinates, linear coefficients and nonlinear parameters (mbx_amd.autograd).
    A step is the forward of both families, the loss and the full backward."""
    import numpy as np
    import torch

    from mbx_amd import autograd as ad
    from mbx_amd import configs

    cfg = configs.C4
    xd, xt = configs.build_c4(cfg)
    g = torch.Generator(device="cpu").manual_seed(cfg["seed"])
    leaves, refs = {}, {}
    for fam, x in (("mbpol_2b", xd), ("mbpol_3b", xt)):
        coef, theta = ad.defaults(fam)
        xx = torch.from_numpy(x).to("cuda")
        with torch.no_grad():
            e, f = ad.energy_forces(xx, coef, theta, fam)
        refs[fam] = (e + 0.1 * torch.randn(e.shape, generator=g, dtype=torch.float64).cuda(),
                     f + 0.1 * torch.randn(f.shape, generator=g, dtype=torch.float64).cuda())
        leaves[fam] = [xx.requires_grad_(), coef.requires_grad_(), theta.requires_grad_()]

    def step():
        for v in leaves.values():
            for t in v:
                t.grad = None
        loss = 0.0
        for fam, (x, coef, theta) in leaves.items():
            e, f = ad.energy_forces(x, coef, theta, fam)
            loss = loss + ad.energy_force_loss(e, f, *refs[fam])
        loss.backward()
        return loss

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # forward alone, for the split
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.no_grad():
        for _ in range(steps):
            for fam, (x, coef, theta) in leaves.items():
                ad.energy_forces(x, coef, theta, fam)
    torch.cuda.synchronize()
    dtf = (time.perf_counter() - t0) / steps
    n = cfg["ndimers"] + cfg["ntrimers"]
    return {"workload": cfg["name"], "dimers": cfg["ndimers"], "trimers": cfg["ntrimers"],
            "ms_per_step": dt * 1e3, "ms_forward": dtf * 1e3, "nmers_per_s": n / dt, "loss": float(loss),
            "step": "forward (E, F) of both families + energy/force MSE loss + backward to coordinates, "
                    "1153 + 1163 coefficients and 2 x 10 nonlinear parameters"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain legs")
    ap.add_argument("--solo", action="store_true",
                    help="profiling: only the stage-timed steps (each kernel alone on the stream), so a "
                         "profiler's per-kernel averages are the launches the roofline is computed from")
    ap.add_argument("--diff-only", action="store_true",
                    help="only run the differentiation leg (4096 dimers + 4096 trimers, loss and backward)")
    ap.add_argument("--pimd-only", action="store_true",
                    help="only run the batched-vs-plain leg of the 32-bead CH4/H2O workload (profiling)")
    ap.add_argument("--schedule", default=None,
                    help="fixed polynomial schedule 'S2,S3' (wavefronts per 64 dimers, trimers) instead of "
                         "autotuning; used for profiles, so the profiled launches are the timed ones")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 512 MiB device-to-device copy 5 times (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X; --gpus %d is not supported (multi-GPU sharding is "
              "design-only in this build)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8, "copies": 5}))
        return

    from mbx_amd import configs, engine

    if args.pimd_only:
        print(json.dumps({"pimd_batched": pimd_leg(args.steps, args.warmup)}))
        return
    if args.diff_only:
        print(json.dumps({"differentiation": differentiation_leg(args.steps, args.warmup)}))
        return

    cfg = configs.C2
    xyz, box, c2, c3 = configs.build(cfg)
    eng = engine.water_engine(cfg["nmon"], box, c2, c3)
    d_xyz = torch.from_numpy(xyz).to("cuda:0")
    torch.cuda.synchronize()

    def step():
        eng.set_coords_device(d_xyz.data_ptr())
        eng.compute()

    if args.schedule:  # fixed schedule, set before the first launch
        s2, s3 = (int(v) for v in args.schedule.split(","))
        eng.set_tuning("split_2b", s2)
        eng.set_tuning("split_3b", s3)
        schedule = {"split_2b": s2, "split_3b": s3}
    for _ in range(args.warmup):
        step()
    eng.synchronize()
    if not args.schedule:
        schedule = eng.autotune(reps=3)  # polynomial-stage schedule for this workload
    if not args.solo:
        eng.set_graph(True)  # steady state: one captured HIP graph per compute
        for _ in range(2):
            step()
        eng.synchronize()
    # throughput: K steps back to back, no host wait in between
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(0 if args.solo else args.steps):
        step()
    eng.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # stage breakdown: K more steps with HIP events around every stage (eager launches)
    eng.set_graph(False)
    eng.set_timing(True)
    poly_ms, stage = [], {}
    for _ in range(args.steps):
        step()
        tm = eng.timings()
        poly_ms.append(tm["poly_3b"])
        for k, v in tm.items():
            stage[k] = stage.get(k, 0.0) + v
    # the polynomial launch without the 3B chain rule (fuse_chain 0: the chain rule as a kernel of
    # its own), for the polynomial's own roofline figure
    unfused = None
    if eng.tuning("fuse_chain") == 1:
        eng.set_tuning("fuse_chain", 0)
        up, uc = [], []
        for _ in range(args.steps):
            step()
            tm = eng.timings()
            up.append(tm["poly_3b"])
            uc.append(tm["chain_3b"])
        eng.set_tuning("fuse_chain", 1)
        unfused = (float(np.mean(up)), float(np.mean(uc)))
    eng.set_timing(False)
    # the same step with the gradient (and energies) copied to the host every step: synchronous,
    # PCIe-inclusive (graph mode again)
    eng.set_graph(True)
    for _ in range(2):
        step()
        eng.results()
    t_d0 = time.perf_counter()
    for _ in range(0 if args.solo else args.steps):
        step()
        eng.results()
    t_d2h = (time.perf_counter() - t_d0) / max(1, args.steps)
    eng.set_graph(False)

    counts = eng.counts()
    n_listed = int(counts[0] + counts[2])
    dt = (t1 - t0) / args.steps if not args.solo else stage["total"] / 1e3 / args.steps
    e2, e3, grad, vir = eng.results()

    flops = json.load(open(os.path.join(ROOT, "mbx_amd", "tables", "flops.json")))
    f3, f2 = flops["mbpol_3b_eg"]["flops"], flops["mbpol_2b_eg"]["flops"]
    paired = eng.tuning("fuse_poly") == 1 and counts[1] > 0 and counts[3] > 0
    chained = paired and eng.tuning("fuse_chain") == 1
    if chained:  # the launch also runs the 3B chain rule: its algorithmic work counts too
        f3 += flops["mbpol_3b_chain"]["flops"]
    t_poly = float(np.mean(poly_ms)) * 1e-3
    achieved = (counts[3] * f3 + (counts[1] * f2 if paired else 0)) / t_poly
    traffic = read_traffic()
    kernel = ("k_poly_eg_pair_chain3<36, 31> (MB-pol 3B and 2B polynomials E+G and the 3B chain rule, one launch)"
              if chained else "k_poly_eg_pair<36, 31> (MB-pol 3B and 2B polynomials E+G, one launch)" if paired
              else "k_poly_eg<36> (MB-pol 3B polynomial E+G)")

    out = {
        "metric": METRIC,
        "value": n_listed / dt,
        "unit": "dimers+trimers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "waters": cfg["nmon"], "box_A": cfg["edge"], "cutoff2b_A": c2,
                   "cutoff3b_A": c3, "listed_dimers": int(counts[0]), "evaluated_dimers": int(counts[1]),
                   "listed_trimers": int(counts[2]), "evaluated_trimers": int(counts[3]),
                   "stage_ms": {k: v / args.steps for k, v in stage.items()}, "e2b": e2, "e3b": e3,
                   "schedule": schedule},
        "roofline": {"bound": "fp64-valu", "kernel": kernel,
                     "achieved": achieved / 1e12, "peak": FP64_VECTOR_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_VECTOR_PEAK, "traffic": traffic,
                     "flops_per_unit": {"trimer": f3, "dimer": f2 if paired else 0},
                     "units_per_launch": {"trimers": int(counts[3]), "dimers": int(counts[1]) if paired else 0},
                     "launch_ms": t_poly * 1e3},
        "cpu_baseline": None,
    }
    # list build (fix, cloud, binning, both n-mer passes): PMC bytes per step over its HIP-event time
    if chained and unfused:
        f3p = flops["mbpol_3b_eg"]["flops"]
        a_u = (counts[3] * f3p + counts[1] * f2) / (unfused[0] * 1e-3)
        out["config"]["polynomial_alone"] = {
            "kernel": "k_poly_eg_pair<36, 31> (fuse_chain 0; the 3B chain rule then runs as k_grad_mbpol3b)",
            "launch_ms": unfused[0], "achieved_TFLOPs": a_u / 1e12, "frac": a_u / FP64_VECTOR_PEAK,
            "chain_3b_ms": unfused[1]}
    out["config"]["with_grad_d2h"] = {
        "ms_per_step": t_d2h * 1e3, "nmers_per_s": n_listed / t_d2h,
        "step": "the headline step plus mbxg_get_results (energies, virial and the 6144 x 3 gradient copied to "
                "host memory, host waits each step)"}
    lb = read_traffic(LIST_TRAFFIC_FILE, "list_build_bytes_per_step:")
    lt = (stage["fix"] + stage["lists_count"] + stage["lists_emit"]) / args.steps * 1e-3
    out["config"]["list_build"] = {"ms": lt * 1e3, "hbm_bytes": lb, "GB_per_s": lb / lt / 1e9 if lb else None,
                                   "frac_of_hbm_peak": lb / lt / HBM_PEAK if lb else None,
                                   "bound": "latency (one 4-wavefront block per query; gathers, LDS sort)"}
    # MB-pol 1B of the same 2048 waters (mbxg_eval_1b: synchronous, host copies of the per-water
    # energies, virial terms and the gradient included)
    eng.eval_1b()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e1b, _, _ = eng.eval_1b()
    t1b = (time.perf_counter() - t0) / args.steps
    out["config"]["one_body"] = {"e1b": e1b, "ms_per_call": t1b * 1e3, "waters_per_s": cfg["nmon"] / t1b,
                                 "call": "mbxg_eval_1b with gradients, results copied to the host"}
    if not args.solo:
        import tempfile

        _, g1b, _ = eng.eval_1b()
        with tempfile.TemporaryDirectory() as tmp:
            di, t_di, g_di = drop_in_leg(xyz, box, c2, c3, args.steps, tmp)
        di["nmers_per_s"] = n_listed / t_di
        # the drop-in gradient (input order = engine order for waters) against the engine's
        # 2B + 3B gradient of the headline step plus its 1B gradient
        di["grad_maxabs_vs_engine"] = float(np.abs(g_di - (grad + g1b)).max())
        out["config"]["drop_in"] = di
    if not args.no_batch:
        out["config"]["batched"] = batched_leg(args.steps, args.warmup)
        out["config"]["pimd_batched"] = pimd_leg(args.steps, args.warmup)
        out["config"]["differentiation"] = differentiation_leg(args.steps, args.warmup)
    if not args.no_cpu:
        cb, r = cpu_baseline(xyz, box, c2, c3, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["config"]["cpu_parity"] = {"e2b_rel": (e2 - r["e2b"]) / r["e2b"], "e3b_rel": (e3 - r["e3b"]) / r["e3b"],
                                       "grad_maxabs": float(np.abs(grad - r["grad"]).max())}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
