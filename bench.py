#!/usr/bin/env python3
"""Throughput of the PMPIC per-timestep parcel-in-cell path on one MI355X.

Workload: the moist thermal on a 128x128x128 grid (2pi x 2pi x pi box, periodic x/y, rigid z)
with 27 parcels per cell on a jittered 3x3x3 lattice (56,623,104 parcels), plain FP64, RK4 with
dt = 0.01, splitting and mixing on.  A step is one full timestep (the integrator group run for
the four RK stages: par2grid, vort2vel, vorticity_tendency, grid2par + RK update, haloswap,
and at stage 4 splitting and mixing) with the parcels resident in HBM.

Prints one JSON line: parcel-steps/s over the timed steps, the roofline of the dominant kernel
(HIP events on the library's stream around every launch in the timed region) and a CPU
baseline (the C restatement of the same path, single-threaded, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "parcel-steps/s FP32-bulk vs FP64 on 1 MI355X at ≤1e-5 vel err; HBM GB/s"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per launch of the two hot kernels (see DESIGN.md "Roofline accounting")
#  cell accumulation: per parcel the bucket entry (4 B) and x,y,z,vol,p,q,r,b,h (72 B);
#    per cell two bucket offsets (8 B) and 40 corner partial sums written (320 B)
#  gather + RK stage: per parcel position (24 B), RK state read/written for the stage, six
#    tendencies written (48 B), direction flag (1 B); per node the six gathered grids (48 B)
RK_STAGE_BYTES = {1: 24 + 24 + 18 * 8 + 48 + 1, 2: 24 + 12 * 8 + 12 * 8 + 48 + 1,
                  3: 24 + 12 * 8 + 12 * 8 + 48 + 1, 4: 24 + 7 * 8 + 13 * 8 + 48 + 1}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="cells per horizontal direction")
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=3, help="parcels per cell per direction")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (the reference kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibrate(mib):
    """hipMemcpy device-to-device of `mib` MiB: the runtime's copy kernel reads and writes
    exactly that many bytes, which calibrates the HBM traffic counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes}))


def cpu_baseline(steps):
    """The oracle (C FP64 restatement, one thread) on the 32^3, 27 parcels/cell thermal."""
    from oracle import oracle
    from pmpic_amd.config import fortran_real
    from pmpic_amd.initial import thermal
    oracle.build()
    lx, lz = fortran_real("6.28318531"), fortran_real("3.14159265")
    ic = thermal(32, 32, 32, lx, lx, lz, 3, seed=0x504D504943000000)
    o = oracle.Oracle(32, 32, 32, 0.0, lx, 0.0, lx, 0.0, lz, 2 * len(ic["x"]), 0.01, 0.01, 0.5)
    o.set_parcels(ic)
    total = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
        total += o.n
    dt = time.perf_counter() - t0
    return {"value": total / dt, "unit": "parcel-steps/s", "cores": 1, "kind": "port",
            "sample": f"C FP64 restatement of the same path, 32x32x32 cells, 27 parcels/cell "
                      f"({len(ic['x'])} parcels), {steps} steps, single thread, {dt:.1f} s"}


def main():
    args = parse()
    if args.copy_calibrate:
        copy_calibrate(args.copy_calibrate)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        print("bench.py: only --gpus 1 is supported (the whole domain is on one MI355X)",
              file=sys.stderr)
        return 2
    from pmpic_amd.config import fortran_real
    from pmpic_amd.initial import thermal
    from pmpic_amd.lib import Context

    n, nz, ppcd = args.n, args.nz, args.ppc
    lx, lz = fortran_real("6.28318531"), fortran_real("3.14159265")
    t_ic = time.perf_counter()
    ic = thermal(n, n, nz, lx, lx, lz, ppcd, seed=0x504D504943000000 + 1)
    npar = len(ic["x"])
    print(f"bench: initial condition {npar} parcels in {time.perf_counter() - t_ic:.1f} s",
          file=sys.stderr, flush=True)
    ctx = Context(n, n, nz, 0.0, lx, 0.0, lx, 0.0, lz, 2 * npar, dtm=0.01, dtmax=0.01,
                  timestep_prefactor=0.5)
    ctx.set_parcels(ic)
    del ic
    for s in range(args.warmup):
        ctx.step(1)
        print(f"bench: warmup step {s + 1} done, {ctx.n} parcels", file=sys.stderr, flush=True)
    ctx.timing_enable(True)
    total = 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        ctx.step(1)  # returns after the library's stream has drained
        total += ctx.n
    t1 = time.perf_counter()
    elapsed = t1 - t0
    print(f"bench: {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    p2g_ms, p2g_calls = ctx.timing_read(0)
    rk_ms, rk_calls = ctx.timing_read(1)
    ncell = n * n * nz
    nnode = n * n * (nz + 1)
    avg_np = total / max(args.steps, 1)
    if p2g_ms >= rk_ms:
        name = "k_cell_accum (par2grid per-cell corner sums)"
        per_launch = avg_np * 76 + ncell * 328
        avg_ms = p2g_ms / max(p2g_calls, 1)
    else:
        name = "k_rk_stage (grid2par of 6 fields + RK4 stage update)"
        per_launch = avg_np * sum(RK_STAGE_BYTES.values()) / 4 + nnode * 48
        avg_ms = rk_ms / max(rk_calls, 1)
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    value = total / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "parcel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered-lattice moist thermal)",
        "config": {"workload": f"moist thermal {n}x{n}x{nz} cells, {ppcd ** 3} parcels/cell "
                               f"({npar} parcels), plain FP64, dt=0.01, split/mix on",
                   "parcels": npar, "grid": [n, n, nz], "form": "plain FP64"},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_ms": avg_ms, "bytes_per_launch": per_launch,
                     "p2g_ms_total": p2g_ms, "p2g_launches": p2g_calls,
                     "rk_ms_total": rk_ms, "rk_launches": rk_calls},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
