"""Parcel dump files in the reference's binary format.

components/writebinary/src/writebinary.F90:165-205 writes, as an unformatted stream file:
time (FP64), xmin xmax ymin ymax zmin zmax (6 x FP64), the parcel count (int64), then the 17
parcel arrays x .. tag (FP64, n each) and the qvalues (none on this path);
components/readfromfile_parcelsetup/src/readfromfile_parcelsetup.F90:217-316 reads the same.
"""
import numpy as np

FIELDS = ["x", "y", "z", "p", "q", "r", "dxdt", "dydt", "dzdt", "dpdt", "dqdt", "drdt",
          "h", "b", "vol", "stretch", "tag"]


def write_parcels(path, time, bounds, fields):
    n = len(fields["x"])
    with open(path, "wb") as fh:
        np.array([time], dtype="<f8").tofile(fh)
        np.asarray(bounds, dtype="<f8").reshape(6).tofile(fh)
        np.array([n], dtype="<i8").tofile(fh)
        for k in FIELDS:
            np.ascontiguousarray(fields.get(k, np.zeros(n)), dtype="<f8").tofile(fh)


def read_parcels(path):
    with open(path, "rb") as fh:
        time = float(np.fromfile(fh, dtype="<f8", count=1)[0])
        bounds = np.fromfile(fh, dtype="<f8", count=6)
        n = int(np.fromfile(fh, dtype="<i8", count=1)[0])
        fields = {k: np.fromfile(fh, dtype="<f8", count=n) for k in FIELDS}
    return time, bounds, fields
