"""A run from a config.mcf file: the grid, the initial parcels, the timestep loop, the dumps and
the timer summary, as the reference's model driver wires its components (monc_driver.F90,
model_core/src/monc.F90:130-235, global_config).

    python -m pmpic_amd.driver --config=path/to/config.mcf [--key=value ...]

* Grid and run options: nx, ny, nz, xmin .. zmax, max_parcels, dtm, dtmax,
  timestep_prefactor, parcel_splitting_enabled, parcel_mixing_enabled (setup_grid.F90:44-104,
  parcelsetup_base.F90:49, rk4_integrator.F90:38-39, vorticity_tendency.F90:96); ``precision``
  (0 plain FP64, 1 FP32-bulk) selects the numeric form.
* Initial parcels, by the enabled set-up component in the reference's initialisation order
  (global_config:48): ``plume_parcelsetup`` (pmpic_amd/plume.py), ``readfromfile_parcelsetup``
  (the multi-file restart of dump ``restart_num``, which also sets the iteration count), or
  ``basic_parcelsetup``, here the seeded jittered-lattice thermal of pmpic_amd/initial.py with
  ``parcels_per_cell_dir`` parcels per cell and direction.
* The loop is control.run_model with writebinary (``writebinary_enabled``) in the start group
  and the timer summary at the end when ``timing_enabled``.
"""
import sys

from . import control, dumpio
from .config import apply_cli, get_real, parse_mcf


def _flag(options, key):
    v = options.get(key, False)
    if isinstance(v, str):
        return v.strip().lower() in (".true.", "true", "t")
    return bool(v)


def context_from_options(options):
    from .lib import Context
    nx, ny, nz = int(options["nx"]), int(options["ny"]), int(options["nz"])
    box = [get_real(options, k) for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")]
    return Context(nx, ny, nz, *box, int(options["max_parcels"]),
                   dtm=get_real(options, "dtm"), dtmax=get_real(options, "dtmax"),
                   timestep_prefactor=get_real(options, "timestep_prefactor"),
                   splitting=_flag(options, "parcel_splitting_enabled"),
                   mixing=_flag(options, "parcel_mixing_enabled"),
                   precision=int(options.get("precision", 0)))


def initial_parcels(options, nx, ny, nz, box, directory=".", log=print):
    """(parcel arrays, model time, iteration count) from the enabled parcel set-up."""
    if _flag(options, "plume_parcelsetup_enabled"):
        from .plume import plume_from_options
        fields, _ = plume_from_options(options, nx, ny, nz, *box, log=log)
        return fields, 0.0, 0
    if _flag(options, "readfromfile_parcelsetup_enabled"):
        num = int(options["restart_num"])
        time, fields = dumpio.read_restart(directory, num, box[0], box[1], box[2], box[3],
                                           log=log)
        return fields, time, num
    from .initial import thermal
    ppcd = int(options.get("parcels_per_cell_dir", 2))
    fields = thermal(nx, ny, nz, box[1] - box[0], box[3] - box[2], box[5] - box[4], ppcd,
                     xmin=box[0], ymin=box[2], zmin=box[4])
    return fields, 0.0, 0


def run(options, directory=".", log=print):
    """Set up and run the model from parsed options; returns (RunState, context)."""
    ctx = context_from_options(options)
    fields, time, iterations = initial_parcels(options, ctx.nx, ctx.ny, ctx.nz, ctx.box,
                                               directory, log)
    ctx.set_parcels(fields)
    ctx.set_time(time, ctx.dtm)
    writer = None
    if _flag(options, "writebinary_enabled"):
        writer = control.WriteBinary(options, directory, log)
    st = control.run_model(ctx, options, log=log, writer=writer, iterations=iterations,
                           timers=_flag(options, "timing_enabled"),
                           parcels_dat=f"{directory}/parcels.dat")
    return st, ctx


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    cfg = [a.split("=", 1)[1] for a in argv if a.startswith("--config=")]
    if not cfg:
        print("usage: python -m pmpic_amd.driver --config=config.mcf [--key=value ...]")
        return 2
    options = apply_cli(parse_mcf(cfg[0]), [a for a in argv if not a.startswith("--config=")])
    st, ctx = run(options)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
