"""Benchmark: beams/s of the collapsed-cone dose engine on one MI355X.

Workload (default): the heterogeneous single beam -- 128^3 voxels of 2.5 mm
with a bone slab (depth 3-5 cm, rho 1.85) and a lung slab (5-12 cm, rho
0.25) plus N(0, 0.02^2) density noise (seed 1234), 6 MV point source at SAD
100 cm, 10 x 10 cm field, 12 x 12 = 144 cones of the varian_6 kernel.
One step = one beam: fused TERMA + support box + collapsed-cone gather, with
densities, aperture raster and cone plan resident in HBM.

Printed JSON line: value = beams/s; plus voxel-cone updates/s, DDA-step
updates/s, per-step latency p50/p99, the max dose error against the FP64
oracle (oracle TERMA -> oracle convolve_c) in units of Dmax, a roofline
object for the gather kernel (12 B per voxel-cone update: FP32 TERMA read +
FP32 dose read-modify-write) timed with HIP events on its own stream, and
the CPU baseline: the FP64 oracle timed on this host's cores for the same
beam.
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
BYTES_PER_VOXEL_CONE = 12


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="oracle threads (default: min(cpus, 16))")
    p.add_argument("--tile", default="auto",
                   help="gather tile TXxTYxTZ or 'auto' (autotune)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="first run one device copy kernel of MIB MiB (the "
                        "known-size reference for PMC byte counters)")
    return p.parse_args()


def traffic_from_profile():
    """HBM bytes per gather launch from the committed rocprofv3 PMC
    summary, if present (line 'traffic_bytes_per_launch: N')."""
    path = os.path.join(ROOT, "profiles", "gather_hetero_pmc.md")
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)",
                          f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        return 2

    from conehead_amd import _lib, autotune, engine as E
    from conehead_amd.device import DeviceArray, Event, Stream
    from conehead_amd.workloads import hetero

    eng = E.Engine()
    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        a = DeviceArray((n,), np.uint8, zero=True)
        b = DeviceArray((n,), np.uint8)
        _lib.check(_lib.lib().ch_memcpy_d2d(b.ptr, a.ptr, n, None))
        _lib.check(_lib.lib().ch_device_sync())
        del a, b
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
                                stream=s, use_cache=False)
    else:
        tile = tuple(int(v) for v in args.tile.split("x"))
        plan.set_tile(*tile)

    def step():
        E.terma([beam], shape, case.origin, case.spacing, stream=s,
                check=False)
        E.convolve(plan, d_t, d_d, d_box, stream=s)

    for _ in range(args.warmup):
        step()
    s.sync()

    # timed region: K back-to-back steps, events at every step boundary and
    # around every gather launch (all on stream s)
    K = args.steps
    ev = [Event() for _ in range(K + 1)]
    g0 = [Event() for _ in range(K)]
    g1 = [Event() for _ in range(K)]
    s.sync()
    t_start = time.perf_counter()
    ev[0].record(s)
    for i in range(K):
        E.terma([beam], shape, case.origin, case.spacing, stream=s,
                check=False)
        g0[i].record(s)
        E.convolve(plan, d_t, d_d, d_box, stream=s)
        g1[i].record(s)
        ev[i + 1].record(s)
    s.sync()
    wall = time.perf_counter() - t_start
    _lib.check(_lib.lib().ch_terma_check(s.handle))
    step_ms = np.array([ev[i].elapsed_ms(ev[i + 1]) for i in range(K)])
    gather_ms = np.array([g0[i].elapsed_ms(g1[i]) for i in range(K)])
    gpu_total_ms = ev[0].elapsed_ms(ev[K])
    ms_per_step = wall * 1e3 / K
    beams_per_s = K / wall
    dose = d_d.download(stream=s)
    terma_gpu = d_t.download(stream=s)

    # CPU baseline and parity: the FP64 oracle on this host for the beam
    result_err = None
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import ref
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        e, w, mu = E.spectrum(st)
        rb = ref.make_beam(src.position, src.v_y, src.transform, src.sad,
                           st, e, w, mu)
        c0 = time.perf_counter()
        out = ref.terma(rb, shape, case.origin, case.spacing, case.densities,
                        blk.block_values, n_threads=threads)
        ref_dose = np.zeros(shape)
        updates = ref.convolve_dirs(out["terma"], ref_dose, case.spacing,
                                    case.cones.dirs, case.cones.rows,
                                    case.cones.cum, n_threads=threads)
        cpu_s = time.perf_counter() - c0
        dmax = float(ref_dose.max())
        result_err = {
            "dose_max_abs_err_over_dmax":
                float(np.abs(dose - ref_dose).max() / dmax),
            "terma_max_abs_err_over_tmax":
                float(np.abs(terma_gpu - out["terma"]).max() /
                      out["terma"].max()),
        }
        cpu = {"value": 1.0 / cpu_s, "unit": "beams/s", "cores": threads,
               "kind": "port",
               "sample": "one full beam of this workload: FP64 TERMA + "
                         "convolve_c scatter over all %d T!=0 voxels x %d "
                         "cones (%d dose += T*k updates), %.2f s" %
                         (int(np.count_nonzero(out["terma"])), n_cones,
                          updates, cpu_s),
               "dda_step_updates": int(updates)}

    g_ms = float(gather_ms.mean())
    achieved = BYTES_PER_VOXEL_CONE * nv * n_cones / (g_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": beams_per_s,
        "unit": "beams/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 DDA alphas)",
        "data": "synthetic (seeded phantom; EDKnrc varian_6 kernel)",
        "config": {
            "workload": "heterogeneous single beam: 128^3 @ 2.5 mm, lung "
                        "and bone slabs + N(0,0.02^2) noise, 10x10 cm field, "
                        "12x12 = 144 cones, 1 beam per step",
            "grid": list(shape), "voxel_mm": 2.5, "cones": n_cones,
            "beams_per_step": 1, "gather_tile": list(tile),
        },
        "voxel_cone_updates_per_s": nv * n_cones * beams_per_s,
        "dda_step_updates_per_s": (cpu["dda_step_updates"] * beams_per_s
                                   if cpu else None),
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                       "p99": float(np.percentile(step_ms, 99))},
        "gpu_ms_per_step": gpu_total_ms / K,
        "gather_ms": g_ms,
        "terma_ms": float((step_ms - gather_ms).mean()),
        "error_vs_fp64_oracle": result_err,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profile(),
        },
        "cpu_baseline": cpu,
        "device": _lib.device_info(),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
