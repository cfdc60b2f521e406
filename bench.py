"""Benchmark: MLUPS and HBM-roofline fraction of the two-component D3Q27 TSLB
step on one MI355X.

Workload: the emulsion configuration -- 512^3 periodic, two components,
200 non-overlapping droplets with radii in [8, 16], surface tension 0.01,
recolouring beta 0.7, linear forcing, FP32 compute and FP32 moment storage.
One step = colour-gradient normals + fused moment-space collide/pull-stream over
the whole lattice, inputs resident in HBM.

Prints one JSON line.  `roofline` is for the dominant (fused) kernel:
algorithmic bytes per launch (88 B per lattice update: 11 FP32 fields read
once and written once) over its mean launch time, measured with HIP events on
the library's stream inside the timed region.  `cpu_baseline` times the CPU
oracle (the reference's population-push algorithm, FP32) on a slab of the same
workload on this host's cores.  Other workloads (--workload): the channel,
the Taylor-Green vortex, the colour-gradient sphere (11 fields, its colour
gradient formed inside the fused pass: 88 B per lattice update) and the
closure LES.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline, 2C D3Q27 512³; closure-LES speedup at fixed error"
HBM_PEAK_GBS = 8000.0
FIELDS_2C = 11


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may use: its CPU affinity, capped by OMP_NUM_THREADS
    when the host sets it (a shared GPU host allots each GPU its share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(n, 1)


def cpu_baseline(n, case, budget_s=15.0, threads=None):
    """Time the FP32 CPU oracle (population push, as the reference) on a slab
    n x n x nzs of the emulsion workload, periodic, on every core this process
    may use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    threads = threads or host_cores()
    po.set_threads(threads)
    nzs = max(4, min(32, n // 16))
    f = {k: np.ascontiguousarray(v[:nzs]) for k, v in case.fields.items()}
    kw = case.lattice
    o = po.Oracle(n, n, nzs, ncomp=2, tau=kw["tau"], sigma=kw["sigma"], sharp_c=kw["sharp_c"],
                  lin_a=kw["lin_force"], periodic=(1, 1, 1), dtype=np.float32)
    from entropicthreadsafe_amd.configs import rand_noseeded_grid
    vel = {}
    for name, amp, l in case.noise:   # the same seeded noise as the GPU run, first nzs planes
        r = rand_noseeded_grid(n, n, nzs, l)
        vel[name] = (f[name] + np.float32(amp) * (np.float32(2) * r - np.float32(1))).astype(np.float32)
    o.init(f["rho"], vel["u"], vel["v"], vel["w"], f["phi"], True)
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter() - t0
    steps = max(2, int(budget_s / max(t1, 1e-3)))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    mlups = n * n * nzs * steps / dt / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(),
            "sample": f"FP32 population-push oracle, emulsion slab {n}x{n}x{nzs} periodic, "
                      f"{steps} steps in {dt:.1f} s on {threads} threads (all cores allotted to "
                      f"this process; the host has {os.cpu_count()})"}


def load_traffic(path):
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=None,
                    help="lattice edge (default 512; 256 for the colour sphere)")
    ap.add_argument("--workload", default="emulsion",
                    choices=("emulsion", "channel", "tgv", "colour", "closure-les"),
                    help="emulsion (2C, the named workload), channel (1C, n x n/2 x n/2), tgv (1C, n^3), "
                         "colour (colour-gradient sphere, 256^3), "
                         "closure-les (learned closure on n/4 against the filtered n^3 run)")
    ap.add_argument("--storage", default="fp32", choices=("fp32", "fp16", "bf16", "fx16"),
                    help="moment storage precision (compute is FP32)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the emulsion line at ~15 %% volume fraction")
    ap.add_argument("--no-les", action="store_true",
                    help="skip the closure-LES figures of the default line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1 (the multi-GPU slab "
              "exchange is a design, not a built path)", file=sys.stderr)
        sys.exit(2)

    from entropicthreadsafe_amd import configs
    from entropicthreadsafe_amd.tslb import Lattice

    if args.workload == "closure-les":
        closure_les(args)
        return
    n = args.n or (256 if args.workload == "colour" else 512)
    t0 = time.time()
    if args.workload == "emulsion":
        case = configs.emulsion(n=n)
    elif args.workload == "channel":
        case = configs.channel(nx=n, ny=n // 2, nz=n // 2)
    elif args.workload == "colour":
        # colour-gradient sphere of radius 40 at 256^3 (sigma 0.02, beta 0.7)
        case = configs.colour_droplet(n=n, radius=40.0 * n / 256, sigma=0.02, beta=0.7)
    else:
        case = configs.tgv(n=n)
    nz, ny, nx = case.shape
    ncomp = case.lattice["ncomp"]
    log(f"bench: {args.workload} {nx}x{ny}x{nz} set up in {time.time() - t0:.1f} s ({case.notes})")
    m = measure(Lattice, case, args.storage, args.steps, args.warmup, copy=True)
    nodes = nx * ny * nz

    mlups = nodes * args.steps / m["wall"] / 1e6
    nfields = 11 if ncomp >= 2 else 10
    es = 4 if args.storage == "fp32" else 2
    bytes_per_launch = 2.0 * es * nfields * nodes     # read + write of the stored moment fields
    t_fused = m["fused_ms"] / 1e3
    achieved = bytes_per_launch / t_fused / 1e9
    t_step = m["fused_ms"] / 1e3 + m["curv_ms"] / 1e3
    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated
    # with the copy kernel), for the default size and FP32 storage only
    traffic = None
    if n == 512 and args.storage == "fp32":
        traffic = load_traffic(os.path.join(ROOT, "profiles", f"fused_{args.workload}_traffic.json"))
    out = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["wall"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32" else f"f32 compute, {args.storage} storage",
        "data": "synthetic (seeded droplets and velocity noise)",
        "config": {"workload": f"{args.workload}-{ncomp}c-{nx}x{ny}x{nz}", "nx": nx, "ny": ny, "nz": nz,
                   "components": ncomp, "storage": args.storage, "notes": case.notes,
                   **{k: v for k, v in case.lattice.items() if k in ("tau", "sigma", "sharp_c", "lin_force", "force")}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": ("stored figure: rocprofv3 FETCH_SIZE (x2 copy-kernel calibration) + "
                                        "WRITE_SIZE of this kernel at this size, "
                                        f"profiles/fused_{args.workload}_traffic.json; not measured in this run")
                     if traffic else None,
                     "kernel": f"k_fused<{ncomp}>", "algorithmic_bytes_per_launch": bytes_per_launch,
                     "mean_launch_ms": round(t_fused * 1e3, 4)},
        "step_roofline_frac": round(bytes_per_launch / t_step / 1e9 / HBM_PEAK_GBS, 4),
        "hbm_stream_copy_gbs": round(m["copy_gbs"], 1),
        "fused_frac_of_copy": round(achieved / m["copy_gbs"], 4),
        "curvature_ms_per_step": round(m["curv_ms"], 4),
    }
    if args.workload == "emulsion" and n == 512 and not args.no_dense:
        # the same workload at the ~15 % volume fraction its description
        # names: 2900 droplets of radii U[8, 16] (200 droplets fill 1.2 %)
        t0 = time.time()
        dense = configs.emulsion(n=n, ndrop=2900)
        log(f"bench: dense emulsion set up in {time.time() - t0:.1f} s ({dense.notes})")
        dsteps = max(20, args.steps // 2)
        d = measure(Lattice, dense, args.storage, dsteps, args.warmup)
        out["emulsion_dense"] = {
            "notes": dense.notes, "steps": dsteps,
            "mlups": round(nodes * dsteps / d["wall"] / 1e6, 2),
            "ms_per_step": round(d["wall"] / dsteps * 1e3, 4),
            "fused_ms": round(d["fused_ms"], 4), "curvature_ms": round(d["curv_ms"], 4),
            "curvature_share": round(d["curv_ms"] / (d["fused_ms"] + d["curv_ms"]), 4),
            "fused_roofline_frac": round(bytes_per_launch / (d["fused_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "step_roofline_frac": round(bytes_per_launch / ((d["fused_ms"] + d["curv_ms"]) / 1e3) / 1e9
                                        / HBM_PEAK_GBS, 4),
        }
    if not args.no_cpu and args.workload == "emulsion" and args.storage == "fp32":
        try:
            out["cpu_baseline"] = cpu_baseline(n, case, args.cpu_budget)
        except Exception as e:  # the baseline is reported, never the product
            log(f"bench: cpu baseline failed: {e}")
            out["cpu_baseline"] = None
    if args.workload == "emulsion" and n == 512 and not args.no_les:
        # the second half of the metric: the closure LES against the plain
        # coarse run and the filtered fine run (entropicthreadsafe_amd/les.py)
        try:
            out["closure_les"] = closure_les_summary(les_run(log))
        except Exception as e:
            log(f"bench: closure LES failed: {e}")
            out["closure_les"] = None
    print(json.dumps(out), flush=True)


def measure(Lattice, case, storage, steps, warmup, copy=False):
    """Time `steps` steps of a case after `warmup`, inputs resident in HBM:
    wall time, mean fused-kernel and curvature-pass launch times (HIP events
    on the library's stream), the copy-kernel rate (copy=True) and the
    mass / phase-mass drift over the timed steps."""
    nz, ny, nx = case.shape
    with Lattice(nx, ny, nz, device=0, storage=storage, **case.lattice) as lat:
        case.apply(lat)
        log(f"bench: state on device ({lat.device_bytes() / 2**30:.2f} GiB)")
        copy_gbs = None
        if copy:
            copy_gbs = lat.stream_copy_gbs()
            log(f"bench: plain copy kernel {copy_gbs:.0f} GB/s")
        lat.step(warmup)
        lat.sync()
        r0 = lat.reduce()
        lat.timing_enable(True)
        lat.sync()
        t0 = time.perf_counter()
        lat.step(steps)
        lat.sync()
        wall = time.perf_counter() - t0
        fused_ms, fused_n = lat.timing_read(0)
        norm_ms, norm_n = lat.timing_read(1)
        lat.timing_enable(False)
        r1 = lat.reduce()
    fused = fused_ms / max(fused_n, 1)
    curv = norm_ms / max(fused_n, 1)       # per step (the pass runs once per step)
    log(f"bench: {steps} steps in {wall:.3f} s; fused {fused:.3f} ms x {fused_n}, "
        f"curvature {norm_ms / max(norm_n, 1):.3f} ms x {norm_n}; mass drift {(r1[0] - r0[0]) / r0[0]:.2e}, "
        f"phi drift {(r1[1] - r0[1]) / max(r0[1], 1e-30):.2e}")
    return dict(wall=wall, fused_ms=fused, curv_ms=curv, copy_gbs=copy_gbs,
                mass_drift=(r1[0] - r0[0]) / r0[0], phase_drift=(r1[1] - r0[1]) / max(r0[1], 1e-30))


def les_run(log_fn, n=512):
    from entropicthreadsafe_amd import les
    return les.run(n=n, log=log_fn)


def closure_les_summary(res):
    return {
        "speedup_closure_vs_fine": round(res["speedup_closure_vs_fine"], 2),
        "closure_spectrum_error": res["closure_spectrum_error"],
        "plain_spectrum_error": res["plain_spectrum_error"],
        "kmax": res["kmax"],
        "fine_seconds": round(res["fine_seconds"], 4),
        "closure_seconds": round(res["closure_seconds"], 4),
        "plain_seconds": round(res["plain_seconds"], 4),
        "workload": res["workload"],
    }


def closure_les(args):
    """The closure-LES line: wall-time ratio of the fine run to the coarse run
    with the learned closure over the same physical horizon, with the
    energy-spectrum errors of the coarse runs (closure and plain) against the
    filtered fine run."""
    from entropicthreadsafe_amd import les
    res = les.run(n=args.n or 512, log=log)
    out = {
        "metric": "closure-LES speedup at fixed error",
        "value": round(res["speedup_closure_vs_fine"], 2),
        "unit": "x (fine wall time / closure coarse wall time, same physical horizon)",
        "n_gpus": 1,
        "steps": res["horizon_fine_steps"],
        "warmup": res["spinup"],
        "ms_per_step": round(res["closure_seconds"] / res["horizon_coarse_steps"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lattice, bf16 closure operands",
        "data": "synthetic (seeded solenoidal field and droplets)",
        "config": {"workload": res["workload"], "factor": res["factor"], "kmax": res["kmax"],
                   "train_steps": res["train_steps"], "batch": res["batch"], "samples": res["samples"]},
        "closure_spectrum_error": res["closure_spectrum_error"],
        "balanced_spectrum_error": res.get("balanced_spectrum_error"),
        "nu_t_balance": res.get("nu_t_balance"),
        "plain_spectrum_error": res["plain_spectrum_error"],
        "fine_seconds": round(res["fine_seconds"], 4),
        "closure_seconds": round(res["closure_seconds"], 4),
        "plain_seconds": round(res["plain_seconds"], 4),
        "train_seconds": round(res["train_seconds"], 3),
        "train_loss": [res["train_loss_initial"], res["train_loss_final"]],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
