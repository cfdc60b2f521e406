"""Long runs of the channel and emulsion workloads: conserved sums and
kinetic energy every `every` steps, in any storage format.

    python -m entropicthreadsafe_amd.longrun channel fp32 bf16 fx16 --steps 20000
    python -m entropicthreadsafe_amd.longrun emulsion fp32 fp16 --steps 20000

The channel is the 512x256x256 Re_tau 180 case and the emulsion the 512^3
200-droplet case (configs.channel, configs.emulsion: the reference's
20k-step channel and emulsion configurations).  One line per report: mass
and phase-mass drift relative to the start, kinetic energy and enstrophy
(FP64 sums of tslb_reduce), whether they are finite, and the wall time.
"""
import argparse
import time

import numpy as np

from entropicthreadsafe_amd import configs
from entropicthreadsafe_amd.tslb import Lattice


def run(workload, storage, steps, every, out=print):
    case = configs.emulsion() if workload == "emulsion" else configs.channel()
    nz, ny, nx = case.shape
    t0 = time.time()
    rows = []
    with Lattice(nx, ny, nz, storage=storage, **case.lattice) as lat:
        case.apply(lat)
        r0 = lat.reduce()
        out(f"{workload} {storage}: {case.notes}; set up in {time.time() - t0:.1f} s")
        t1 = time.time()
        done = 0
        while done < steps:
            n = min(every, steps - done)
            lat.step(n)
            done += n
            r = lat.reduce()
            dm = (r[0] - r0[0]) / r0[0]
            dp = (r[1] - r0[1]) / r0[1] if r0[1] != 0 else 0.0
            rows.append((done, dm, dp, r[2], r[6]))
            out(f"  step {done:6d}: mass {dm:+.2e} phase {dp:+.2e} KE {r[2]:.4e} enstrophy {r[6]:.4e} "
                f"finite {bool(np.all(np.isfinite(r)))} ({time.time() - t1:.1f} s)")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("workload", choices=("channel", "emulsion"))
    ap.add_argument("storage", nargs="+", choices=("fp32", "fp16", "bf16", "fx16"))
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--every", type=int, default=2000)
    a = ap.parse_args(argv)
    for st in a.storage:
        run(a.workload, st, a.steps, a.every, out=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main()
