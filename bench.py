#!/usr/bin/env python3
"""Throughput of the PMPIC per-timestep parcel-in-cell path on one MI355X.

Workload (the precision form's configuration, the one the metric is quoted on): the moist
thermal on a 256x256x128 grid (2pi x 2pi x pi box, periodic x/y, rigid z) with 8 parcels per
cell on a jittered 2x2x2 lattice (67,108,864 parcels), RK4 with dt = 0.01, splitting and mixing
on.  A step is one full timestep (the integrator group run for the four RK stages: par2grid,
vort2vel, vorticity_tendency, grid2par + RK update, haloswap, and at stage 4 splitting and
mixing) with the parcels resident in HBM.

Both numeric forms run on the same initial condition: the FP32-bulk form (positions as cell
index + FP32 offsets, FP32 parcel arithmetic; `value`) and the plain FP64 form.  The velocity
error of the FP32-bulk form against the FP64 form is measured over --err-steps steps of both,
stepped side by side from that same state: relative L2 error of the u, v, w grids (the bar,
1e-5), the maximum-norm error beside it, and the step at which the two forms' parcel counts
first differ.  Prints one JSON line: parcel-steps/s of each form, the roofline of each
form's dominant kernel and of the other parcel<->grid kernel (HIP events taken by every launch
in the timed region), the HBM ceilings measured in the same run (device copy, STREAM triad) and
a CPU baseline: the threaded C FP64 restatement of the same path on all host cores for one step
of the same workload.
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

# algorithmic bytes per launch of the two hot kernels
#  cell accumulation: per parcel the bucket entry (4 B) and x,y,z,vol,p,q,r,b,h (72 B);
#    per cell two bucket offsets (8 B) and 40 corner partial sums written (320 B)
#  gather + RK stage: per parcel position (24 B), RK state read/written for the stage, six
#    tendencies written (48 B), direction flag (1 B); per node the six gathered grids (48 B)
RK_STAGE_BYTES = {1: 24 + 24 + 18 * 8 + 48 + 1, 2: 24 + 12 * 8 + 12 * 8 + 48 + 1,
                  3: 24 + 12 * 8 + 12 * 8 + 48 + 1, 4: 24 + 7 * 8 + 13 * 8 + 48 + 1}
# FP32-bulk form: cell accumulation reads per parcel its slot entry, offsets, vol, p,q,r,b,h
# (40 B; the z level comes from the cell) and per cell its count and overflow offset (8 B);
# it writes 40 FP32 corner partials per cell (160 B), or, when nx is a multiple of 16 (tiles
# of 16 cells in one x row), the partials combined along x: 4 corner pairs x 17 node positions
# x 5 values per 16 cells (85 B per cell).  The RK stages move cell (4 B) + offsets (12 B) for
# the position and the FP32 RK state (see bulk32.hip)
ACC_BYTES = {0: (76, 328), 1: (40, 168)}
ACC_BYTES_XROWS = (40, 8 + 85)
RK32_STAGE_BYTES = {1: 28 + 80, 2: 68 + 52, 3: 68 + 52, 4: 60 + 56}
# on grids of 2^21 cells or more, RK stages 1-3 also bin the cells they write: a 4-byte slot
# entry per parcel (the count atomics, one per run of a cell's parcels, left out)
RK32_BIN_BYTES = 4
GATHER_NODE_BYTES = {0: 48, 1: 24}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="cells per horizontal direction")
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--ppc", type=int, default=2, help="parcels per cell per direction")
    ap.add_argument("--n2", type=float, default=1.0,
                    help="buoyancy frequency squared of the stable layer (4 for the large run)")
    ap.add_argument("--cap-factor", type=float, default=2.0,
                    help="parcel capacity as a multiple of the initial parcel count")
    ap.add_argument("--cpu-steps", type=int, default=1,
                    help="steps of the workload the CPU baseline times")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forms", default="fp32,fp64",
                    help="numeric forms to time: fp32 (FP32-bulk), fp64 (plain), or both")
    ap.add_argument("--err-steps", type=int, default=100,
                    help="steps of the FP32-bulk vs FP64 velocity-error run on the workload "
                         "(0 = skip)")
    ap.add_argument("--autotune", type=int, default=0, metavar="K",
                    help="with --ensemble: pick the members' launch configurations first, "
                         "timing the ensemble stepped together (ensemble.tune_together, K timed "
                         "steps per candidate; untimed model steps)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="M",
                    help="instead of the single thermal: M FP32-bulk thermals on grids cycling "
                         "32^3, 48^3, 64^3, 96^3 (8 or 27 parcels/cell, radius U(0.5,1), "
                         "amplitude U(0.5,2), own seeds), timed stepped together and one after "
                         "another")
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


def copy_bandwidth(mib=1024, reps=5):
    """The practical HBM ceiling on this GPU: the best of `reps` device-to-device hipMemcpy of
    `mib` MiB (the runtime's copy kernel), timed with HIP events; bytes = read + write."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    best = None
    for _ in range(reps + 1):  # the first copy is a warm-up
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        best = ms.value if best is None else min(best, ms.value)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return {"GBps": 2 * nbytes / (best * 1e-3) / 1e9, "bytes": 2 * nbytes, "ms": best,
            "what": f"hipMemcpy device-to-device of {mib} MiB, best of {reps}"}


def cpu_baseline(ic, n, nz, steps):
    """The oracle (C FP64 restatement of the same path, OpenMP over every host core, results
    independent of the thread count) timed for `steps` whole steps of the bench's own workload
    from its initial condition."""
    from oracle import oracle
    oracle.build()
    lx, lz = ic["_lx"], ic["_lz"]
    npar = len(ic["x"])
    o = oracle.Oracle(n, n, nz, 0.0, lx, 0.0, lx, 0.0, lz, npar + npar // 4 + 1024, 0.01, 0.01,
                      0.5)
    o.set_parcels({k: v for k, v in ic.items() if not k.startswith("_")})
    threads = oracle.max_threads()
    total = 0
    t0 = time.perf_counter()
    for s in range(steps):
        o.step()
        total += o.n
        print(f"bench: cpu baseline step {s + 1} done", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    del o
    return {"value": total / dt, "unit": "parcel-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(),
            "sample": f"C FP64 restatement of the same path (oracle/, OpenMP, {threads} threads), "
                      f"{steps} step(s) of the bench workload: {n}x{n}x{nz} cells, {npar} "
                      f"parcels, {dt:.1f} s"}


def traffic_file(n, nz, ppcd):
    return os.path.join(ROOT, "profiles",
                        f"hbm_traffic_thermal{n}x{n}x{nz}_{ppcd ** 3}ppc.txt")


def measured_traffic(form, family, n, nz, ppcd):
    """HBM bytes per launch of a kernel family from the committed rocprofv3 counter summary of
    this workload (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration); None when
    no summary covers the workload."""
    path = traffic_file(n, nz, ppcd)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 7 and parts[0] == "traffic" and parts[1] == form and parts[2] == family:
            return float(parts[6])
    return None


def time_form(precision, ic, args, n, nz):
    """Warm up, then time args.steps whole steps of one numeric form; returns the figures."""
    from pmpic_amd.lib import Context
    lx, lz = ic["_lx"], ic["_lz"]
    npar = len(ic["x"])
    ctx = Context(n, n, nz, 0.0, lx, 0.0, lx, 0.0, lz, int(args.cap_factor * npar), dtm=0.01,
                  dtmax=0.01, timestep_prefactor=0.5, precision=precision)
    ctx.set_parcels(ic)
    sums0 = ctx.conservation_sums()
    tag = "fp32-bulk" if precision else "fp64"
    for s in range(args.warmup):
        ctx.step(1)
        print(f"bench[{tag}]: warmup step {s + 1} done, {ctx.n} parcels", file=sys.stderr,
              flush=True)
    # the timed steps carry only the two hot kernels' timers (taken by their dispatches)
    ctx.timing_enable(True, kernels_only=True)
    total = 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        ctx.step(1)  # returns after the library's stream has drained
        total += ctx.n
    elapsed = time.perf_counter() - t0
    print(f"bench[{tag}]: {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    p2g_ms, p2g_calls = ctx.timing_read(0)
    rk_ms, rk_calls = ctx.timing_read(1)
    # parcel count and conservation drift over the warm-up and timed steps
    sums1 = ctx.conservation_sums()
    drift = {"steps": args.warmup + args.steps, "parcels_start": npar, "parcels_end": ctx.n,
             "rel_drift": {k: (sums1[i] - sums0[i]) / sums0[i]
                           for i, k in enumerate(("vol", "vol_b", "vol_h"))}}
    # the component timers (reference region names) from two further, untimed steps
    ctx.timing_enable(True)
    ctx.step(2)
    timers = {k: {"ms_per_step": v[0] / 2, "count": v[1]} for k, v in ctx.timer_table().items()}
    ctx.close()
    ncell = n * n * nz
    nnode = n * n * (nz + 1)
    avg_np = total / max(args.steps, 1)
    form = "fp32" if precision else "fp64"
    xrows = precision == 1 and n % 16 == 0  # the FP32 accumulation's x-combined partials
    acc_bytes = ACC_BYTES_XROWS if xrows else ACC_BYTES[precision]
    acc = {
        "kernel": ("k32_cell_accum (FP32 par2grid cell sums, corner partials combined along x)"
                   if xrows else
                   "k32_cell_accum (FP32 par2grid per-cell corner sums)" if precision else
                   "k_cell_accum (par2grid per-cell corner sums)"),
        "family": ("k_cell_accum_lds<0>" if not precision else
                   "k32_cell_accum<3>" if xrows else "k32_cell_accum<0>"),
        "bytes": avg_np * acc_bytes[0] + ncell * acc_bytes[1],
        "ms": p2g_ms, "calls": p2g_calls,
    }
    table = dict(RK32_STAGE_BYTES if precision else RK_STAGE_BYTES)
    fenv = os.environ.get("PMPIC_FUSED_BINNING")
    fused = precision == 1 and (fenv == "1" if fenv else ncell >= (1 << 21))
    if fused:
        for s_ in (1, 2, 3):
            table[s_] += RK32_BIN_BYTES
    rks = {
        "kernel": ("k32_rk_stage (FP32 grid2par of 6 fields + RK4 stage update"
                   + (", cell binning in stages 1-3)" if fused else ")") if precision else
                   "k_rk_stage (grid2par of 6 fields + RK4 stage update)"),
        "family": "k32_rk_stage" if precision else "k_rk_stage",
        "bytes": avg_np * sum(table.values()) / 4 + nnode * GATHER_NODE_BYTES[precision],
        "ms": rk_ms, "calls": rk_calls,
    }

    def roof(k):
        avg_ms = k["ms"] / max(k["calls"], 1)
        achieved = k["bytes"] / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": k["kernel"], "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(form, k["family"], n, nz, args.ppc),
                "avg_launch_ms": avg_ms, "bytes_per_launch": k["bytes"],
                "launches": k["calls"], "ms_total": k["ms"]}

    dominant, other = (acc, rks) if p2g_ms >= rk_ms else (rks, acc)
    return {
        "value": total / elapsed,
        "ms_per_step": 1e3 * elapsed / max(args.steps, 1),
        # the dominant kernel, and the other half of the parcel<->grid pair beside it
        "roofline": roof(dominant),
        "roofline_other": roof(other),
        "timers": timers,
        "conservation": drift,
    }


def velocity_error(args, ic, n, nz):
    """Both forms from the bench's initial condition, stepped side by side for args.err_steps
    steps.  The FP32-bulk form's velocity error relative to the FP64 form over the u, v, w grids,
    in the L2 norm (||u32 - u64|| / ||u64||, the bar) and in the maximum norm (max |u32 - u64| /
    max |u64|), at checkpoints, with the parcel counts of both forms and the first step at which
    they differ (a split or mixing decision taken differently at its threshold)."""
    import numpy as np
    from pmpic_amd.lib import Context
    lx, lz = ic["_lx"], ic["_lz"]
    ctx = {}
    for precision in (1, 0):
        c = Context(n, n, nz, 0.0, lx, 0.0, lx, 0.0, lz, 2 * len(ic["x"]), dtm=0.01,
                    dtmax=0.01, timestep_prefactor=0.5, precision=precision)
        c.set_parcels(ic)
        ctx[precision] = c

    def errors():
        g32 = [ctx[1].grid(k)[2:-2, 2:-2, :] for k in ("u", "v", "w")]
        g64 = [ctx[0].grid(k)[2:-2, 2:-2, :] for k in ("u", "v", "w")]
        l2 = np.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(g32, g64)) /
                     sum(float(np.sum(b ** 2)) for b in g64))
        mx = (max(float(np.max(np.abs(a - b))) for a, b in zip(g32, g64)) /
              max(float(np.max(np.abs(b))) for b in g64))
        return l2, mx

    history = []
    first_diff = None
    checkpoints = {1, 10, 25, 50, 75, args.err_steps}
    for s in range(1, args.err_steps + 1):
        ctx[1].step(1)
        ctx[0].step(1)
        differ = ctx[1].n != ctx[0].n
        if differ and first_diff is None:
            first_diff = s
            l2, mx = errors()
            history.append([s, l2, mx, ctx[1].n, ctx[0].n])
        if s in checkpoints:
            l2, mx = errors()
            history.append([s, l2, mx, ctx[1].n, ctx[0].n])
            print(f"bench: error run at step {s}: L2 {l2:.3e}, max {mx:.3e}, parcels "
                  f"{ctx[1].n} / {ctx[0].n}", file=sys.stderr, flush=True)
    l2, mx = history[-1][1], history[-1][2]
    out = {"rel_velocity_error_l2": l2, "rel_velocity_error_max": mx, "steps": args.err_steps,
           "grid": [n, n, nz], "parcels_fp32": ctx[1].n, "parcels_fp64": ctx[0].n,
           "first_step_counts_differ": first_diff,
           "history": {"columns": ["step", "l2", "max", "parcels_fp32", "parcels_fp64"],
                       "rows": history},
           "ok": bool(l2 <= 1e-5)}
    for c in ctx.values():
        c.close()
    return out


def ensemble(args):
    """Batched-ensemble figure: M FP32-bulk members (pmpic_amd/ensemble.py: grids cycling 32^3,
    48^3, 64^3, 96^3 cells, 8 or 27 parcels per cell, own radius, amplitude and seed) stepped
    concurrently (pmpic_step_concurrent) and the same members stepped one after another, from the
    same initial states."""
    from pmpic_amd import ensemble as ens
    from pmpic_amd.lib import step_concurrent
    specs = ens.member_specs(args.ensemble)
    ics = [ens.member_parcels(s) for s in specs]
    total = sum(len(ic["x"]) for ic in ics)
    print(f"bench: ensemble of {len(ics)} members, {total} parcels", file=sys.stderr, flush=True)
    out = {}
    tuned = {}
    if args.autotune:
        # tuned on copies of the members (the tuning steps are model steps), so the timed
        # members start from the same states as untuned ones
        ctxs = [ens.member_context(s, ic) for s, ic in zip(specs, ics)]
        tuned = ens.tune_together(ctxs, args.autotune)
        for c in ctxs:
            c.close()
        print(f"bench: ensemble tuned: {tuned}", file=sys.stderr, flush=True)
    for mode in ("concurrent", "sequential"):
        ctxs = [ens.member_context(s, ic) for s, ic in zip(specs, ics)]
        if args.autotune:
            for c in ctxs:
                for cls, cfg in tuned["by_slot_capacity"].items():
                    if c.launch_config()["slot_capacity"] == cls:
                        c.set_launch_config(0, **cfg)
        if mode == "concurrent":
            step_concurrent(ctxs, args.warmup)
        else:
            for c in ctxs:
                c.step(args.warmup)
        n0 = sum(c.n for c in ctxs)
        t0 = time.perf_counter()
        if mode == "concurrent":
            step_concurrent(ctxs, args.steps)
        else:
            for c in ctxs:
                c.step(args.steps)
        el = time.perf_counter() - t0
        n1 = sum(c.n for c in ctxs)
        out[mode] = 0.5 * (n0 + n1) * args.steps / el
        print(f"bench: ensemble {mode}: {args.steps} steps in {el:.3f} s", file=sys.stderr,
              flush=True)
        for c in ctxs:
            c.close()
    print(json.dumps({
        "metric": "ensemble parcel-steps/s (FP32-bulk)", "value": out["concurrent"],
        "unit": "parcel-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded moist thermals)",
        "config": {"workload": f"{len(ics)} thermals on grids cycling 32^3..96^3, 8 or 27 "
                               f"parcels/cell",
                   "members": [[s[0], s[1] ** 3] for s in specs], "parcels": total},
        "sequential_value": out["sequential"],
        "autotuned": tuned or None,
        "concurrency_speedup": out["concurrent"] / out["sequential"]}))
    return 0


def main():
    args = parse()
    if args.ensemble:
        return ensemble(args)
    if args.copy_calibrate:
        copy_calibrate(args.copy_calibrate)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "the whole domain runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        print("bench.py: only --gpus 1 is supported (the whole domain is on one MI355X)",
              file=sys.stderr)
        return 2
    from pmpic_amd.config import fortran_real
    from pmpic_amd.initial import thermal

    n, nz, ppcd = args.n, args.nz, args.ppc
    lx, lz = fortran_real("6.28318531"), fortran_real("3.14159265")
    t_ic = time.perf_counter()
    ic = thermal(n, n, nz, lx, lx, lz, ppcd, seed=0x504D504943000000 + 1, n2=args.n2)
    ic["_lx"], ic["_lz"] = lx, lz
    npar = len(ic["x"])
    print(f"bench: initial condition {npar} parcels in {time.perf_counter() - t_ic:.1f} s",
          file=sys.stderr, flush=True)
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    for f in forms:
        res[f] = time_form(1 if f == "fp32" else 0, ic, args, n, nz)
    head = res.get("fp32") or res["fp64"]
    form = "FP32-bulk" if "fp32" in res else "plain FP64"
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "parcel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if "fp32" in res else "f64",
        "data": "synthetic (seeded jittered-lattice moist thermal)",
        "config": {"workload": f"moist thermal {n}x{n}x{nz} cells, {ppcd ** 3} parcels/cell "
                               f"({npar} parcels), N2={args.n2:g}, {form}, dt=0.01, "
                               f"split/mix on",
                   "parcels": npar, "grid": [n, n, nz], "n2": args.n2, "form": form},
        "roofline": head["roofline"],
        "roofline_other": head["roofline_other"],
        "timers": head["timers"],
        "conservation": head["conservation"],
    }
    if "fp64" in res:
        out["fp64"] = {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms_per_step"],
                       "roofline": res["fp64"]["roofline"],
                       "roofline_other": res["fp64"]["roofline_other"],
                       "timers": res["fp64"]["timers"],
                       "conservation": res["fp64"]["conservation"]}
    if "fp32" in res and "fp64" in res:
        out["fp32_speedup_vs_fp64"] = res["fp32"]["value"] / res["fp64"]["value"]
    # the measured ceilings beside the 8 TB/s peak: the runtime's device copy and a STREAM triad
    out["hbm_copy"] = copy_bandwidth()
    from pmpic_amd.lib import hbm_triad
    tms, tgbps = hbm_triad(1 << 30, 5)
    out["hbm_triad"] = {"GBps": tgbps, "ms": tms, "bytes": 3 << 30,
                        "what": "STREAM triad a = b + s*c over three 1 GiB FP64 arrays "
                                "(pmpic_hbm_triad), best of 5"}
    rl = [out["roofline"], out["roofline_other"]]
    if "fp64" in out:
        rl += [out["fp64"]["roofline"], out["fp64"]["roofline_other"]]
    for r in rl:
        r["frac_of_copy"] = r["achieved"] / out["hbm_copy"]["GBps"]
        r["frac_of_triad"] = r["achieved"] / tgbps
    if args.err_steps > 0 and "fp32" in res:
        out["fp32_error"] = velocity_error(args, ic, n, nz)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ic, n, nz, args.cpu_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
