"""Benchmark: beams/s of the collapsed-cone dose engine on one MI355X.

Default workload (`value`): the heterogeneous single beam -- 128^3 voxels of
2.5 mm with a bone slab (depth 3-5 cm, rho 1.85) and a lung slab (5-12 cm,
rho 0.25) plus N(0, 0.02^2) density noise (seed 1234), 6 MV point source at
SAD 100 cm, 10 x 10 cm field, 12 x 12 = 144 cones of the varian_6 kernel.
One step = one beam: fused TERMA, then the cone convolution -- the FFT form
(five passes) or the direct stencil gather (support box, padding, gather),
whichever is faster (both are timed) -- with densities, aperture, cone plan
and kernel spectrum resident in HBM.  `value` replays the step as a HIP graph
K times (its dose checked bitwise against the eager step's); an eager loop
with events at every step gives latency, the per-call split and the kernel
times.

The same JSON line carries
  roofline      the step's dominant kernel over its launch time, measured
                with HIP events the library attaches to the kernels on the
                stream they run on: TERMA (its sort pass and brick kernel;
                8 B per voxel) or the convolution (the FFT form's
                convolution pass: 8 B per complex64 bin read or written, or
                the gather: 12 B per voxel-cone update), whichever takes
                longer; `rooflines` has both and `step_roofline` the whole
                step's algorithmic bytes over the replayed step time;
  cpu_baseline  the FP64 oracle (TERMA + convolve_c) timed on this host's
                cores for the same beam;
  batched       64 MLC beams at random gantry angles on the same phantom,
                96 cones: the beams' TERMA sharing launches, then their
                convolutions sharing launches -- eager, and captured once
                as one graph from device-resident tables and replayed --
                against the engine's single-beam graph step replayed over
                the same 64 beams (speedup_vs_single_beam_graph);
  plain         the plain single-beam GPU path (per-stage synchronisation
                and host round trips, scatter convolution with atomics) on
                beams of the batched workload, and the batched/plain ratio;
  uneven        32 phantoms of differing grids, voxel sizes, beam counts and
                cone sets as one batch (the gather's tile and the FFT form's
                lines per workgroup autotuned per plan), with the error of
                one beam against the oracle;
  serve         Poisson request streams drawn from the uneven workload,
                lambda swept 10-500 req/s through the request loop (p50/p99
                latency, beams/s), its capacity at 2000 req/s, and one
                served dose against the oracle.
voxel_cone_updates_per_s and dda_step_updates_per_s are the metric's work
units at the delivered beam rate -- the dense count Nv x C per beam and the
reference's count of dose += T * k executions; the FFT form reaches the same
doses without performing them one by one, so for it they are dense
equivalents (updates_are_dense_equivalent).
--workload batched  the 64-beam MLC batch alone (one step = one batch).
--workload uneven   the uneven batch alone, with both convolutions.
--workload serve    the request loop alone, 64 requests per lambda.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("beams/sec + voxel·cone updates/s at ≤0.1% Dmax err; "
          "p50/p99 latency")
HBM_PEAK_GBS = 8000.0
MI355X_CLOCK_GHZ = 2.4      # peak engine clock
BYTES_PER_VOXEL_CONE = 12


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", default="hetero",
                   choices=["hetero", "batched", "uneven", "serve"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the batched and plain sections")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="oracle threads (default: min(cpus, 16))")
    p.add_argument("--tile", default="auto",
                   help="gather tile TXxTYxTZ or 'auto' (autotune)")
    p.add_argument("--plain-beams", type=int, default=4)
    p.add_argument("--method", default="auto",
                   choices=["auto", "gather", "fft"],
                   help="convolution of the headline and batched numbers: "
                        "the direct stencil gather, the FFT form, or the "
                        "faster of the two (both are always timed)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="first run one device copy kernel of MIB MiB (the "
                        "known-size reference for PMC byte counters)")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profile(name="gather_hetero_pmc.md"):
    """HBM bytes per launch of a kernel from a committed rocprofv3 PMC
    summary under profiles/, if present (line 'traffic_bytes_per_launch:
    N')."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)",
                          f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def count_from_profile(name, key):
    """A figure 'key: N' from a committed profile summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            m = re.search(r"%s:\s*([0-9.eE+]+)" % re.escape(key), f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def valu_issue(kernel_ms, profile, key):
    """VALU issue time of a launch from its profiled wave64 VALU
    instruction count: one instruction every 4 cycles on each of the
    device's SIMDs (4 per CU) at the peak engine clock."""
    n = count_from_profile(profile, key)
    if n is None or not kernel_ms:
        return None
    from conehead_amd import _lib
    simds = 4 * (_lib.device_info().get("n_cu") or 256)
    clock_ghz = MI355X_CLOCK_GHZ
    issue_ms = n * 4 / (simds * clock_ghz * 1e9) * 1e3
    return {"valu_instructions": n, "issue_ms": issue_ms,
            "frac": issue_ms / kernel_ms,
            "what": "wave64 VALU instructions per launch (profiles/%s) at "
                    "one per 4 cycles per SIMD, %d SIMDs, %.2f GHz, over "
                    "the launch time" % (profile, simds, clock_ghz)}


def cpu_threads(args):
    return args.cpu_threads or min(os.cpu_count() or 1, 16)


def oracle_beam(case, b, threads):
    """FP64 oracle dose of beam b of a case; returns (dose, terma, updates,
    seconds)."""
    from conehead_amd import engine as E
    from oracle import ref
    src, blk, st = case.beams[b]
    e, w, mu = E.spectrum(st)
    rb = ref.make_beam(src.position, src.v_y, src.transform, src.sad, st, e,
                       w, mu)
    c0 = time.perf_counter()
    out = ref.terma(rb, case.shape, case.origin, case.spacing,
                    case.densities, blk.block_values, n_threads=threads)
    dose = np.zeros(case.shape)
    updates = ref.convolve_dirs(out["terma"], dose, case.spacing,
                                case.cones.dirs, case.cones.rows,
                                case.cones.cum, n_threads=threads)
    return dose, out["terma"], updates, time.perf_counter() - c0


def oracle_beam_one_thread(case, b, terma, stride=16):
    """The oracle on one thread, as the reference runs (GIL held, single
    threaded): the full TERMA, then the convolution over every stride-th
    T!=0 source voxel, scaled to the whole beam by the ratio of dose += T*k
    executions.  Returns (seconds per beam, sample seconds, description)."""
    from conehead_amd import engine as E
    from oracle import ref
    src, blk, st = case.beams[b]
    e, w, mu = E.spectrum(st)
    rb = ref.make_beam(src.position, src.v_y, src.transform, src.sad, st, e,
                       w, mu)
    c0 = time.perf_counter()
    ref.terma(rb, case.shape, case.origin, case.spacing, case.densities,
              blk.block_values, n_threads=1)
    t_terma = time.perf_counter() - c0
    nz = np.flatnonzero(terma)
    pick = nz[::stride]
    ts = np.zeros(terma.shape)
    ts.flat[pick] = terma.flat[pick]
    dose = np.zeros(case.shape)
    c0 = time.perf_counter()
    u_sample = ref.convolve_dirs(ts, dose, case.spacing, case.cones.dirs,
                                 case.cones.rows, case.cones.cum, n_threads=1)
    t_conv = time.perf_counter() - c0
    return t_terma, t_conv, u_sample, len(pick), len(nz)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(),
            "cpus_usable": len(os.sched_getaffinity(0)),
            "cpu_model": model}


def conv_half_forms(n, frame):
    """The convolution pass's template flag H for n voxels on a frame of
    `frame` points: bit 0 when its forward transform's first stage skips the
    padding rows, bit 1 when its inverse transform's last stage skips the
    cropped outputs (both need frame >= 2 n and an even radix in that stage).
    The radices are ordered as the library orders them: 16, 8, 4, 2, then 3
    and 5, with the last even radix moved to the end when two or more are
    even and an odd one would come last."""
    facs, m = [], frame
    for r in (16, 8, 4, 2, 3, 5):
        while m % r == 0:
            facs.append(r)
            m //= r
    even = 0
    while even < len(facs) and facs[even] % 2 == 0:
        even += 1
    if 2 <= even < len(facs):
        facs.append(facs.pop(even - 1))
    if len(facs) < 2 or 2 * n > frame:
        return 0
    return int(facs[0] % 2 == 0) | (int(facs[-1] % 2 == 0) << 1)


def err_over_dmax(got, ref):
    return float(np.abs(got - ref).max() / ref.max())


# --------------------------------------------------------------- c2 ------
def run_hetero(args, eng):
    from conehead_amd import _lib, autotune, engine as E
    from conehead_amd.device import (DeviceArray, Event, Stream,
                                     arm_launch_events)
    from conehead_amd.workloads import hetero

    case = hetero()
    shape, nv = case.shape, case.n_voxels
    n_cones = case.cones.n_cones
    src, blk, st = case.beams[0]
    s = Stream()
    d_rho = DeviceArray.from_host(case.densities, stream=s)
    d_t = DeviceArray(shape, np.float32)
    d_d = DeviceArray(shape, np.float32)
    d_box = DeviceArray((6,), np.int32)
    d_r = eng.aperture(blk, s)
    beam = E.beam_struct(src, st, d_r, d_rho, d_t, d_box)
    plan = eng.plan(shape, case.spacing, case.cones)

    E.terma([beam], shape, case.origin, case.spacing, stream=s)
    box = d_box.download(stream=s)
    if args.tile == "auto":
        tile, _ = autotune.tune(plan, box=tuple(int(b) for b in box),
                                stream=s, use_cache=False, terma=d_t)
    else:
        tile = tuple(int(v) for v in args.tile.split("x"))
        plan.set_tile(*tile)

    # the FFT form needs no support box: its TERMA skips computing one
    beam_fft = E.beam_struct(src, st, d_r, d_rho, d_t, None)

    def conv_fn(method):
        if method == "gather":
            return lambda: E.convolve(plan, d_t, d_d, d_box, stream=s)
        plan.spectrum(s)
        return lambda: E.convolve_fft(plan, d_t, d_d, stream=s)

    def timed(method):
        """K back-to-back steps (TERMA + convolution) with events at every
        step boundary (ev), around every convolve call (g0, g1) and around
        the method's dominant kernel, recorded by the library (k0, k1):
        the gather, or the FFT form's convolution pass; then the same step
        captured once as a HIP graph and replayed K times."""
        conv = conv_fn(method)
        bm = beam_fft if method == "fft" else beam
        for _ in range(args.warmup):
            E.terma([bm], shape, case.origin, case.spacing, stream=s,
                    check=False)
            conv()
        s.sync()
        K = args.steps
        ev = [Event() for _ in range(K + 1)]
        g0 = [Event() for _ in range(K)]
        g1 = [Event() for _ in range(K)]
        k0 = [Event() for _ in range(K)]
        k1 = [Event() for _ in range(K)]
        q0 = [Event() for _ in range(K)]
        q1 = [Event() for _ in range(K)]
        s.sync()
        t_start = time.perf_counter()
        ev[0].record(s)
        for i in range(K):
            arm_launch_events(q0[i], q1[i])
            E.terma([bm], shape, case.origin, case.spacing, stream=s,
                    check=False)
            g0[i].record(s)
            arm_launch_events(k0[i], k1[i])
            conv()
            g1[i].record(s)
            ev[i + 1].record(s)
        s.sync()
        wall = time.perf_counter() - t_start
        _lib.check(_lib.lib().ch_terma_check(s.handle))
        r = {"wall": wall,
             "step_ms": np.array([ev[i].elapsed_ms(ev[i + 1])
                                  for i in range(K)]),
             "call_ms": np.array([g0[i].elapsed_ms(g1[i])
                                  for i in range(K)]),
             "kernel_ms": np.array([k0[i].elapsed_ms(k1[i])
                                    for i in range(K)]),
             "terma_kernels_ms": np.array([q0[i].elapsed_ms(q1[i])
                                           for i in range(K)]),
             "gpu_total_ms": ev[0].elapsed_ms(ev[K]),
             "dose": d_d.download(stream=s)}
        with E.Graph.capture(s) as cap:
            E.terma([bm], shape, case.origin, case.spacing, stream=s,
                    check=False)
            conv()
        graph = cap.graph
        graph.launch(s)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(K):
            graph.launch(s)
        s.sync()
        r["graph_wall"] = time.perf_counter() - t0
        r["graph_ms"] = r["graph_wall"] * 1e3 / K
        graph.close()
        # the replayed step must give the eager step's dose, bit for bit
        if not np.array_equal(d_d.download(stream=s), r["dose"]):
            raise RuntimeError("%s: graph replay differs from the eager "
                               "step" % method)
        return r

    K = args.steps
    runs = {}
    for method in ("gather", "fft"):
        log("hetero: %s" % method)
        runs[method] = timed(method)
    terma_gpu = d_t.download(stream=s)
    frame, pitch, _ = plan.spectrum(s)
    kz = frame[2] // 2 + 1
    # algorithmic bytes per launch of each method's dominant kernel
    kernel_bytes = {
        # one FP32 T read + one FP32 dose read-modify-write per voxel-cone
        "gather": BYTES_PER_VOXEL_CONE * nv * n_cones,
        # convolution pass: the x < X half-spectrum slabs read and written
        # once (complex64) and K^ read once
        "fft": 8 * kz * frame[1] * (2 * shape[0] + frame[0]),
    }
    traffic_file = {"gather": "gather_hetero_pmc.md",
                    "fft": "spectral_hetero_pmc.md"}
    # the FFT form's five passes over the half spectra (complex64, 8 B):
    # z forward (T read, X x Y x Kz written), y forward (read, X x Fy x Kz
    # written), x convolution (read, K^ read, write), y inverse, z inverse
    hs = 8 * shape[0] * shape[1] * kz
    full = 8 * shape[0] * frame[1] * kz
    step_conv_bytes = {
        "gather": BYTES_PER_VOXEL_CONE * nv * n_cones,
        "fft": (4 * nv + hs) + (hs + full) + kernel_bytes["fft"] +
               (full + hs) + (hs + 4 * nv)}
    lines = 16                      # the convolution pass's line group
    half_forms = conv_half_forms(shape[0], frame[0])
    while lines > 2 and frame[0] * lines > 4096:
        lines //= 2
    kernel_name = {"gather": "gather_single (direct stencil gather)",
                   "fft": "col_kernel<%d, 2, %d> (kColConv: x transforms, "
                          "times K^, inverse x transforms)"
                          % (lines, half_forms)}
    # TERMA: the sort pass and the brick kernel, bracketed by events the
    # library attaches to them; 8 B per voxel (a 4 B density read and a 4 B
    # TERMA write)
    terma_bytes = 8 * nv
    for method, r in runs.items():
        k_ms = float(r["kernel_ms"].mean())
        achieved = kernel_bytes[method] / (k_ms * 1e-3) / 1e9
        # headline: the step replayed as a HIP graph, no events inside
        r["beams_per_s"] = K / r["graph_wall"]
        r["eager_beams_per_s"] = K / r["wall"]
        conv = {"bound": "hbm", "kernel": kernel_name[method],
                "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "algorithmic_bytes": kernel_bytes[method],
                "kernel_ms": k_ms,
                "traffic": traffic_from_profile(traffic_file[method])}
        t_ms = float(r["terma_kernels_ms"].mean())
        ta = terma_bytes / (t_ms * 1e-3) / 1e9
        terma = {"bound": "hbm",
                 "kernel": "terma_sort_kernel + terma_kernel (TERMA: hit "
                           "test, FP64 depth walk, fluence, spectrum sum)",
                 "achieved": ta, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": ta / HBM_PEAK_GBS, "algorithmic_bytes": terma_bytes,
                 "kernel_ms": t_ms,
                 "traffic": traffic_from_profile("terma_hetero_pmc.md"),
                 # what bounds it: the FP64 depth walks' instruction issue
                 "valu_issue": valu_issue(
                     t_ms, "terma_conv_hetero_sq_counters.md",
                     "terma_valu_instructions_per_launch")}
        r["rooflines"] = {"terma": terma, "convolution": conv}
        # the step's dominant kernel (the longer of the two)
        r["roofline"] = terma if t_ms >= k_ms else conv
        # the whole step: every kernel's algorithmic bytes over the step
        step_bytes = terma_bytes + step_conv_bytes[method]
        sa = step_bytes / (r["graph_ms"] * 1e-3) / 1e9
        r["step_roofline"] = {
            "bound": "hbm", "achieved": sa, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": sa / HBM_PEAK_GBS,
            "algorithmic_bytes": step_bytes,
            "what": "TERMA's 8 B per voxel plus the convolution's "
                    "algorithmic bytes (%s) over the graph-replayed step" %
                    ("the five FFT passes' complex64 reads and writes and "
                     "K^" if method == "fft" else
                     "12 B per voxel-cone update")}
    if args.method == "auto":
        head = max(runs, key=lambda m: runs[m]["beams_per_s"])
    else:
        head = args.method
    hr = runs[head]
    wall, step_ms, call_ms = hr["wall"], hr["step_ms"], hr["call_ms"]
    gpu_total_ms, graph_ms, dose = (hr["gpu_total_ms"], hr["graph_ms"],
                                    hr["dose"])

    # the convolve_c drop-in on host float64 buffers: PCIe-inclusive rate
    from conehead_amd.convolve_c import convolve_c
    from conehead_amd.kernel2 import merged_kernel
    from conehead_amd.workloads import base_kernel
    kern = merged_kernel(base_kernel(), len(case.cones.phis))
    t64 = terma_gpu.astype(np.float64)
    d64 = np.zeros(shape)
    convolve_c(t64, d64, case.spacing, case.cones.thetas, case.cones.phis,
               kern)
    reps = 5
    c0 = time.perf_counter()
    for _ in range(reps):
        convolve_c(t64, d64, case.spacing, case.cones.thetas,
                   case.cones.phis, kern)
    dropin_ms = (time.perf_counter() - c0) * 1e3 / reps

    # physics mode on the same beam: beam-aligned cones, kernel looked up
    # at the radiological distance (EDKnrc kernel, 12 x 12 cones)
    from conehead_amd.cones import ConeSet, beam_frame
    from conehead_amd.kernel import EdkSet, merge_polar
    poly = EdkSet.load().poly(st["energy_weights"])
    phis, cum = merge_polar(poly.angles, poly.kernel_cum,
                            len(case.cones.phis))
    pcs = ConeSet(case.cones.thetas, phis, cum, semantics="physics",
                  frame=beam_frame(src))
    pplan = eng.plan(shape, case.spacing, pcs)
    d_p = DeviceArray(shape, np.float32)
    E.convolve_rad(pplan, d_t, d_rho, d_p, d_box, stream=s)
    p0, p1 = Event(), Event()
    s.sync()
    p0.record(s)
    for _ in range(reps):
        E.convolve_rad(pplan, d_t, d_rho, d_p, d_box, stream=s)
    p1.record(s)
    s.sync()
    physics_ms = p0.elapsed_ms(p1) / reps

    result_err, cpu = None, None
    if not args.no_cpu_baseline:
        threads = cpu_threads(args)
        log("cpu baseline: oracle on %d threads" % threads)
        ref_dose, ref_t, updates, cpu_s = oracle_beam(case, 0, threads)
        result_err = {
            "dose_max_abs_err_over_dmax": err_over_dmax(dose, ref_dose),
            "terma_max_abs_err_over_tmax": err_over_dmax(terma_gpu, ref_t),
        }
        for r in runs.values():
            r["dose_max_abs_err_over_dmax"] = err_over_dmax(r["dose"],
                                                            ref_dose)
        cpu = {"value": 1.0 / cpu_s, "unit": "beams/s", "cores": threads,
               "kind": "port",
               "sample": "one full beam of this workload: FP64 TERMA + "
                         "convolve_c scatter over all %d T!=0 voxels x %d "
                         "cones (%d dose += T*k updates), %.2f s" %
                         (int(np.count_nonzero(ref_t)), n_cones, updates,
                          cpu_s),
               "dda_step_updates": int(updates)}
        log("cpu baseline: oracle on 1 thread, sampled")
        t_terma, t_conv, u_s, n_pick, n_src = oracle_beam_one_thread(
            case, 0, ref_t)
        per_beam = t_terma + t_conv * updates / u_s
        cpu["one_thread"] = {
            "value": 1.0 / per_beam, "unit": "beams/s", "cores": 1,
            "kind": "port",
            "sample": "full FP64 TERMA (%.2f s) + convolve_c scatter over "
                      "every 16th T!=0 source voxel (%d of %d, %d updates, "
                      "%.2f s), scaled to the beam's %d updates" %
                      (t_terma, n_pick, n_src, u_s, t_conv, updates)}

    beams_per_s = hr["beams_per_s"]
    methods = {}
    for m, r in runs.items():
        methods[m] = {
            "beams_per_s": r["beams_per_s"],
            "ms_per_step": r["graph_ms"],
            "eager_beams_per_s": r["eager_beams_per_s"],
            "eager_ms_per_step": r["wall"] * 1e3 / K,
            "kernel_ms": float(r["kernel_ms"].mean()),
            "convolve_call_ms": float(r["call_ms"].mean()),
            "graph_ms_per_step": r["graph_ms"],
            "terma_kernels_ms": float(r["terma_kernels_ms"].mean()),
            "dose_max_abs_err_over_dmax":
                r.get("dose_max_abs_err_over_dmax"),
            "roofline": r["roofline"],
            "rooflines": r["rooflines"],
            "step_roofline": r["step_roofline"]}
    methods["fft"]["frame"] = list(frame)
    return {
        "value": beams_per_s,
        "ms_per_step": graph_ms,
        "config": {
            "workload": "heterogeneous single beam: 128^3 @ 2.5 mm, lung "
                        "and bone slabs + N(0,0.02^2) noise, 10x10 cm field, "
                        "12x12 = 144 cones, 1 beam per step",
            "grid": list(shape), "voxel_mm": 2.5, "cones": n_cones,
            "beams_per_step": 1, "gather_tile": list(tile),
            "method": head,
            "launch": "the step (TERMA + convolution) captured once as a "
                      "HIP graph and replayed K times; its dose is checked "
                      "bitwise against the eager step's",
        },
        # work units of the metric at the delivered beam rate: the dense
        # count Nv x C per beam and the reference's own count of dose +=
        # T * k executions; the FFT form reaches the same doses without
        # performing either one by one, so both are dense-equivalent rates
        "voxel_cone_updates_per_s": nv * n_cones * beams_per_s,
        "dda_step_updates_per_s": (cpu["dda_step_updates"] * beams_per_s
                                   if cpu else None),
        "updates_are_dense_equivalent": head == "fft",
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                       "p99": float(np.percentile(step_ms, 99))},
        "gpu_ms_per_step": gpu_total_ms / K,
        "graph_ms_per_step": graph_ms,
        "eager_ms_per_step": wall * 1e3 / K,
        "eager_beams_per_s": K / wall,
        "kernel_ms": methods[head]["kernel_ms"],
        "convolve_call_ms": float(call_ms.mean()),
        "terma_ms": float((step_ms - call_ms).mean()),
        "terma_kernels_ms": methods[head]["terma_kernels_ms"],
        "methods": methods,
        "error_vs_fp64_oracle": result_err,
        "dropin_convolve_c_ms": dropin_ms,
        "physics_gather_ms": physics_ms,
        "roofline": runs[head]["roofline"],
        "rooflines": runs[head]["rooflines"],
        "step_roofline": runs[head]["step_roofline"],
        "cpu_baseline": cpu,
    }


# --------------------------------------------------------------- c3 ------
def time_batches(be, reqs, s, steps):
    """Median wall and GPU ms of `steps` launches of one batch (after a
    warm-up launch, which also builds plan spectra for the FFT form); the
    first timed batch's doses."""
    from conehead_amd.device import Event
    be.release(be.collect(be.launch(reqs, s), download=False))
    times, gpu, first = [], [], None
    for i in range(steps):
        t0 = time.perf_counter()
        e0 = Event().record(s)
        h = be.launch(reqs, s)
        h.event.sync()
        times.append((time.perf_counter() - t0) * 1e3)
        gpu.append(e0.elapsed_ms(h.event))
        if i == 0:
            first = be.collect(h)
        else:
            be.release(be.collect(h, download=False))
    return float(np.median(times)), float(np.median(gpu)), first


def pick(args, per):
    if args.method != "auto":
        return args.method
    return max(per, key=lambda m: per[m]["beams_per_s"])


def run_batched(args, eng, steps=5):
    from conehead_amd.batch import BatchEngine, DoseRequest
    from conehead_amd.device import Stream
    from conehead_amd.workloads import batched

    case = batched()
    req = DoseRequest.from_case(case, key="batched-phantom")
    s = Stream()
    nb = len(case.beams)
    per, outs, tile = {}, {}, None
    from conehead_amd import autotune
    plan = eng.plan(case.shape, case.spacing, case.cones)
    fft_lines, _ = autotune.tune_fft(plan, n_beams=nb, stream=s,
                                     use_cache=False)
    for method in ("gather", "fft"):
        log("batched: %s" % method)
        be = BatchEngine(eng, method=method)
        if method == "gather":
            tile = tune_requests(eng, be, [req], s)
        else:
            chunk, _ = autotune.tune_pipeline(be, [req], stream=s)
        ms, gms, out = time_batches(be, [req], s, steps)
        outs[method] = out[0]
        per[method] = {"ms_per_batch": ms, "gpu_ms_per_batch": gms,
                       "beams_per_s": nb / (ms * 1e-3)}
    # the gather with interleaved beam groups where boxes allow
    be1 = BatchEngine(eng, group_beams=True)
    ms1, _, _ = time_batches(be1, [req], s, 2)
    per["gather"]["ms_per_batch_with_beam_groups"] = ms1
    # the FFT form captured once and replayed (device-resident tables)
    log("batched: fft captured")
    gms, gout = time_batch_graph(eng, req, steps)
    if not np.array_equal(gout, outs["fft"]):
        raise RuntimeError("captured batch differs from the eager batch")
    per["fft_captured"] = {"ms_per_batch": gms, "gpu_ms_per_batch": gms,
                           "beams_per_s": nb / (gms * 1e-3)}
    outs["fft_captured"] = gout
    # the engine's single-beam graph step on the same 64 beams
    log("batched: single-beam graph steps")
    sms = time_single_beam_graphs(eng, case, outs["fft"], steps)
    if args.method in ("fft", "gather"):
        per = {m: v for m, v in per.items() if m.startswith(args.method)}
    head = pick(argparse.Namespace(method="auto"), per)
    ms = per[head]["ms_per_batch"]
    r = {"workload": "64 MLC beams (gantry U[0,360), seed 2024, 60 leaf "
                     "pairs x 5 mm, apertures 2-20 cm) on the 128^3 "
                     "heterogeneous phantom, 96 cones: TERMA of the beams "
                     "in shared launches, then one launch per convolution "
                     "kernel for all 64",
         "method": head, "gather_tile": list(list(tile.values())[0]),
         "fft_lines_per_workgroup": list(fft_lines),
         "fft_terma_chunk": chunk,
         "ms_per_batch": ms,
         "gpu_ms_per_batch": per[head]["gpu_ms_per_batch"],
         "beams_per_s": nb / (ms * 1e-3),
         "dense_equivalent_voxel_cone_updates_per_s":
             case.n_voxels * case.cones.n_cones * nb / (ms * 1e-3),
         "single_beam_graph_ms_per_beam": sms,
         "ms_per_beam": ms / nb,
         "speedup_vs_single_beam_graph": sms / (ms / nb),
         "methods": per}
    if not args.no_cpu_baseline:
        ref, _, _, cpu_s = oracle_beam(case, 0, cpu_threads(args))
        for m in per:
            per[m]["error_vs_fp64_oracle_beam0"] = err_over_dmax(outs[m][0],
                                                                 ref)
        r["error_vs_fp64_oracle_beam0"] = \
            per[head]["error_vs_fp64_oracle_beam0"]
    return r, case


def time_batch_graph(eng, req, steps):
    """Median GPU ms of `steps` replays of the batch captured as one graph
    (BatchGraph: beam set TERMA + convolution set FFT passes), timed with
    events on the graph's stream; the doses of the last replay."""
    from conehead_amd.batch import BatchGraph
    from conehead_amd.device import Event
    bg = BatchGraph([req], eng)
    try:
        bg.launch()
        bg.stream.sync()
        times = []
        for _ in range(steps):
            e0 = Event().record(bg.stream)
            e1 = bg.launch()
            e1.sync()
            times.append(e0.elapsed_ms(e1))
        return float(np.median(times)), bg.results()[0]
    finally:
        bg.close()


def time_single_beam_graphs(eng, case, want, steps):
    """The engine's single-beam step (TERMA, zero, FFT-form convolution)
    captured once per beam and replayed beam after beam: median GPU ms per
    beam over `steps` passes of all beams; each beam's dose must equal the
    batch's (`want`) bit for bit."""
    from conehead_amd import engine as E
    from conehead_amd.device import DeviceArray, Event, Stream
    s = Stream()
    shape = case.shape
    rho = DeviceArray.from_host(case.densities, stream=s)
    plan = eng.plan(shape, case.spacing, case.cones)
    ts = [DeviceArray(shape) for _ in case.beams]
    ds = [DeviceArray(shape) for _ in case.beams]
    beams = [E.beam_struct(src, st, eng.aperture(blk, s), rho, t, None)
             for (src, blk, st), t in zip(case.beams, ts)]
    # one eager step sizes the stream's TERMA scratch and FFT work grids
    E.terma([beams[0]], shape, case.origin, case.spacing, stream=s)
    E.convolve_fft(plan, ts[0], ds[0], stream=s)
    graphs = []
    try:
        for b, t, d in zip(beams, ts, ds):
            with E.Graph.capture(s) as cap:
                E.terma([b], shape, case.origin, case.spacing, stream=s,
                        check=False)
                d.zero(s)
                E.convolve_fft(plan, t, d, accumulate=True, stream=s)
            graphs.append(cap.graph)
        times = []
        for _ in range(steps + 1):
            e0 = Event().record(s)
            for g in graphs:
                g.launch(s)
            e1 = Event().record(s)
            e1.sync()
            times.append(e0.elapsed_ms(e1) / len(graphs))
        for i, d in enumerate(ds):
            if not np.array_equal(d.download(stream=s), want[i]):
                raise RuntimeError("single-beam graph step differs from the "
                                   "batch for beam %d" % i)
        return float(np.median(times[1:]))
    finally:
        for g in graphs:
            g.close()
        s.sync()


def run_plain(args, case, n_beams):
    from conehead_amd.plain import plain_case
    plain_case(case, 0)                                   # warm-up
    t0 = time.perf_counter()
    doses = [plain_case(case, b)[0] for b in range(n_beams)]
    dt = time.perf_counter() - t0
    return {"workload": "beams 0..%d of the batched workload, one at a "
                        "time through the plain path" % (n_beams - 1),
            "beams": n_beams, "beams_per_s": n_beams / dt,
            "ms_per_beam": dt * 1e3 / n_beams}, doses


# --------------------------------------------------------------- c4 ------
def uneven_requests(per_beam=True):
    from conehead_amd.batch import DoseRequest
    from conehead_amd.workloads import uneven
    cases = uneven()
    return cases, [DoseRequest.from_case(c, per_beam, key=c.name)
                   for c in cases]


def tune_requests(eng, be, reqs, s):
    """Autotune the tile of every distinct plan on its first beam's TERMA,
    with the batched gather the requests will run in."""
    from conehead_amd import autotune
    from conehead_amd import engine as E
    from conehead_amd.device import DeviceArray
    seen = {}
    for r in reqs:
        plan = eng.plan(r.shape, r.spacing, r.cones)
        if plan.handle in seen:
            continue
        d_rho, _ = be.densities(r, s)
        t = DeviceArray(r.shape)
        box = DeviceArray((8,), np.int32)
        src, blk, st = r.beams[0]
        E.terma([E.beam_struct(src, st, eng.aperture(blk, s), d_rho, t,
                               box)], r.shape, r.origin, r.spacing, stream=s)
        b = box.download(stream=s)[:6]
        seen[plan.handle] = autotune.tune(plan, box=tuple(int(v) for v in b),
                                          stream=s, use_cache=False,
                                          terma=t, batch=True)[0]
    return seen


def tune_fft_requests(eng, reqs, s):
    """Autotune the FFT lines per workgroup of every distinct plan, timed on
    as many beams as the requests hold for it (one launch per pass takes
    them all).  Returns the number of plans whose lines changed."""
    from conehead_amd import autotune
    plans = {}
    for r in reqs:
        p = eng.plan(r.shape, r.spacing, r.cones)
        plans.setdefault(p.handle, [p, 0])[1] += len(r.beams)
    changed = 0
    for p, nb in plans.values():
        p.set_fft_lines((0,) * 5)
        before = p.fft_lines()
        lines, _ = autotune.tune_fft(p, n_beams=nb, stream=s,
                                     use_cache=False)
        changed += tuple(lines) != tuple(before)
    return changed


def run_uneven(args, eng, methods=("gather", "fft"), cases_reqs=None):
    """The uneven batch (32 phantoms) as one batch with each method; returns
    (line, the oracle dose of the smallest case's first beam and its index
    or None)."""
    from conehead_amd.batch import BatchEngine
    from conehead_amd.device import Stream
    log("uneven: generating phantoms")
    cases, reqs = cases_reqs or uneven_requests()
    s = Stream()
    nb = sum(len(c.beams) for c in cases)
    upd = sum(c.voxel_cone_updates() * len(c.beams) for c in cases)
    steps = max(1, min(args.steps, 5))
    per, outs, tiles = {}, {}, {}
    fft_tuned = None
    for method in methods:
        log("uneven: %s" % method)
        be = BatchEngine(eng, method=method)
        if method == "gather":
            tiles = tune_requests(eng, be, reqs, s)
        else:
            fft_tuned = tune_fft_requests(eng, reqs, s)
        ms, gms, outs[method] = time_batches(be, reqs, s, steps)
        per[method] = {"ms_per_batch": ms, "gpu_ms_per_batch": gms,
                       "beams_per_s": nb / (ms * 1e-3)}
    head = pick(args, per)
    ms = per[head]["ms_per_batch"]
    r = {"metric": METRIC, "value": nb / (ms * 1e-3), "unit": "beams/s",
         "n_gpus": 1, "steps": steps, "warmup": 1,
         "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": "f32",
         "precision": "FP32 arithmetic; FP64 DDA alphas and TERMA "
                      "geometry",
         "data": "synthetic (seeded phantoms; EDKnrc varian_6 kernel)",
         "config": {"workload": "uneven batch: 32 phantoms (seed 7), dims "
                                "64-200 x 64-200 x 64-160, 2-5 mm voxels, "
                                "lung/bone ellipsoids, 1-9 beams, 48-144 "
                                "cones, one batch",
                    "beams": nb, "method": head,
                    "distinct_tiles": len(set(tiles.values())),
                    "fft_plans_retuned": fft_tuned},
         "voxel_cone_updates_per_s": upd / (ms * 1e-3),
         "methods": per}
    check = None
    if not args.no_cpu_baseline:
        i = int(np.argmin([c.n_voxels * c.cones.n_cones for c in cases]))
        ref, _, _, _ = oracle_beam(cases[i], 0, cpu_threads(args))
        for m in per:
            per[m]["error_vs_fp64_oracle"] = err_over_dmax(outs[m][i][0],
                                                           ref)
        r["error_vs_fp64_oracle"] = {"case": cases[i].name, "beam": 0,
                                     "dose_max_abs_err_over_dmax":
                                         per[head]["error_vs_fp64_oracle"]}
        check = (i, ref)
    return r, check


# --------------------------------------------------------------- c5 ------
def run_serve(args, eng, n_req=64, cases_reqs=None, check=None):
    """Poisson request streams from the uneven workload through the
    request loop, lambda swept 10-500 req/s, then its capacity at 2000
    req/s.  check = (pool index, oracle dose of that request's first beam):
    that request rides first in the lambda = 500 stream and its served
    dose is compared with the oracle."""
    from conehead_amd.batch import BatchEngine
    from conehead_amd.device import Stream
    from conehead_amd.serve import DoseServer, consume, poisson_arrivals
    log("serve: generating phantoms")
    cases, pool = cases_reqs or uneven_requests()
    method = "fft" if args.method == "auto" else args.method
    be = BatchEngine(eng, method=method)
    s = Stream()
    if method == "gather":
        tune_requests(eng, be, pool, s)
    be.release(be.collect(be.launch(pool, s), download=False))  # warm-up
    rng = np.random.default_rng(99)
    sweep = []
    served_err = None
    for lam in (10, 20, 50, 100, 200, 500):
        # the fast streams get three times the requests (still under half
        # a second of arrivals), so their percentiles rest on more samples
        n = n_req if lam < 200 else 3 * n_req
        reqs = [pool[int(i)] for i in rng.integers(0, len(pool), n)]
        if check is not None and lam == 500:
            reqs[0] = pool[check[0]]
        arr = poisson_arrivals(float(lam), n, seed=99 + lam)
        first = []
        # each result is taken as it is done and dropped, as a client would
        with DoseServer(be, max_inflight=3, max_batch_jobs=16) as srv:
            consume(srv, reqs, arr,
                    take=lambda i, r: first.append(r[0]) if i == 0 else None)
        if check is not None and lam == 500:
            served_err = err_over_dmax(first[0], check[1])
        del first
        st = srv.stats()
        st["lambda_req_per_s"] = lam
        sweep.append(st)
        log("serve lambda=%d: p50 %.1f ms p99 %.1f ms %.0f beams/s" %
            (lam, st["latency_ms_p50"], st["latency_ms_p99"],
             st["beams_per_s"]))
    best = max(sweep, key=lambda d: d["beams_per_s"])
    # beyond the sweep: an arrival rate the server cannot keep up with, so
    # beams/s is its capacity, with the doses delivered to the host and
    # with them left in HBM (consumers on the GPU)
    lam, n_sat = 2000, 256      # a longer burst: steadier than 64 requests
    saturation = {}
    for download in (True, False):
        reqs = [pool[int(i)] for i in rng.integers(0, len(pool), n_sat)]
        arr = poisson_arrivals(float(lam), n_sat, seed=99 + lam)
        with DoseServer(be, max_inflight=3, max_batch_jobs=16,
                        download=download) as srv:
            consume(srv, reqs, arr,
                    take=None if download else lambda i, r: be.release([r]))
        st = srv.stats()
        st["lambda_req_per_s"] = lam
        saturation["host" if download else "device_resident"] = st
        log("serve lambda=%d, doses %s: %.0f beams/s" %
            (lam, "to the host" if download else "left in HBM",
             st["beams_per_s"]))
    err = None
    if served_err is not None:
        err = {"case": cases[check[0]].name, "beam": 0, "lambda": 500,
               "dose_max_abs_err_over_dmax": served_err}
    return {"metric": METRIC, "value": best["beams_per_s"],
            "error_vs_fp64_oracle": err,
            "unit": "beams/s", "n_gpus": 1, "steps": n_req * len(sweep),
            "warmup": 1, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f32",
            "precision": "FP32 arithmetic; FP64 DDA alphas and TERMA "
                         "geometry",
            "data": "synthetic (seeded phantoms; EDKnrc varian_6 kernel)",
            "config": {"workload": "serving: Poisson arrivals (seed 99), "
                                   "%d requests per lambda (3x that from "
                                   "200 req/s) drawn from the "
                                   "uneven batch, resident densities, HBM "
                                   "budget admission, 3 stream slots of "
                                   "up to 16 jobs" %
                                   n_req,
                       "method": method},
            "sweep": sweep, "saturation": saturation}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        return 2

    from conehead_amd import _lib
    from conehead_amd import engine as E
    from conehead_amd.device import DeviceArray

    eng = E.Engine()
    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        a = DeviceArray((n,), np.uint8, zero=True)
        b = DeviceArray((n,), np.uint8)
        _lib.check(_lib.lib().ch_memcpy_d2d(b.ptr, a.ptr, n, None))
        _lib.check(_lib.lib().ch_device_sync())
        del a, b

    if args.workload == "uneven":
        line, _ = run_uneven(args, eng)
    elif args.workload == "batched":
        bt, _ = run_batched(args, eng, steps=max(args.steps, 1))
        line = {"metric": METRIC, "value": bt["beams_per_s"],
                "unit": "beams/s", "n_gpus": 1, "steps": max(args.steps, 1),
                "warmup": args.warmup, "ms_per_step": bt["ms_per_batch"],
                "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32",
                "data": "synthetic (seeded phantom and MLC apertures; "
                        "EDKnrc varian_6 kernel)",
                "config": {"workload": bt.pop("workload"),
                           "method": bt["method"]}}
        line.update(bt)
    elif args.workload == "serve":
        line = run_serve(args, eng)
    else:
        log("hetero single beam")
        h = run_hetero(args, eng)
        line = {"metric": METRIC, "value": h.pop("value"),
                "unit": "beams/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": h.pop("ms_per_step"),
                "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f32",
                "precision": "FP32 arithmetic; FP64 DDA alphas and TERMA "
                             "geometry",
                "data": "synthetic (seeded phantom; EDKnrc varian_6 kernel)",
                "config": h.pop("config")}
        line.update(h)
        if not args.no_extras:
            log("batched 64 beams")
            bt, case = run_batched(args, eng)
            log("plain path")
            pl, _ = run_plain(args, case, args.plain_beams)
            pl["batched_over_plain"] = bt["beams_per_s"] / pl["beams_per_s"]
            line["batched"] = bt
            line["plain"] = pl
            # the uneven batch and the request loop on its requests, with
            # the headline convolution
            cr = uneven_requests()
            m = line["config"]["method"]
            un, check = run_uneven(args, eng, methods=(m,), cases_reqs=cr)
            line["uneven"] = {k: un[k] for k in (
                "config", "ms_per_step", "value", "unit",
                "voxel_cone_updates_per_s", "error_vs_fp64_oracle")
                if k in un}
            sv = run_serve(args, eng, n_req=32, cases_reqs=cr, check=check)
            line["serve"] = {k: sv[k] for k in (
                "config", "sweep", "saturation", "error_vs_fp64_oracle")}
    line["device"] = _lib.device_info()
    line["host"] = host_info()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
