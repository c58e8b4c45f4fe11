"""config.mcf front end with the reference's value semantics.

model_core/src/configuration/configurationfileparser.F90:40-166, 279-329 and
model_core/src/utils/conversions.F90:111-116, 445-453:

* ``key=value`` lines, ``#``/``!`` comments, ``global_configuration=<file>`` include (the included
  file is read first, so the including file overrides it), ``key+=value`` appends to an array,
  comma-separated values make an array (``key``, with elements addressed by index), and
  ``--key=value`` command-line overrides win over both files.
* Type inference order: logical (``.true.``/``.false.``), integer, real, string.
* Reals are read as default-real (FP32) from the first 10 characters (``(f10.0)``; a string with
  an exponent uses ``(es30.10)``), then widened with ``dnint(dble(x)*1e8)/1e8``.
"""
import os
import re

import numpy as np


def fortran_real(text):
    """Value of a config real as the reference stores it (FP32 read, 1e-8 rounding to FP64)."""
    s = text.strip()
    if "e" in s.lower():
        v32 = np.float32(float(s))
    else:
        field = s[:10]
        v32 = np.float32(float(field)) if field.strip() else np.float32(0.0)
    d = float(np.float64(v32))
    return float(np.rint(d * 100000000.0) / 100000000.0)


_INT_RE = re.compile(r"^[+-]?\d+$")
_REAL_RE = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdD][+-]?\d+)?$")


def infer(value):
    v = value.strip()
    low = v.lower()
    if low in (".true.", ".false.", "true", "false"):
        return low in (".true.", "true")
    if _INT_RE.match(v):
        return int(v)
    if _REAL_RE.match(v):
        return fortran_real(v.replace("d", "e").replace("D", "e"))
    if len(v) >= 2 and v[0] == v[-1] and v[0] in "\"'":
        return v[1:-1]
    return v


def _strip_comment(v):
    for ch in "#!":
        pos = v.find(ch)
        if pos >= 0:
            v = v[:pos]
    return v


def parse_mcf(path, options=None):
    """Parse an .mcf file (and its global_configuration include) into a dict."""
    options = {} if options is None else options
    base = os.path.dirname(os.path.abspath(path))
    lines = open(path).read().splitlines()
    local = []
    for raw in lines:
        line = raw.strip()
        if not line or line[0] in "#!":
            continue
        if "=" not in line:
            continue
        key, val = line.split("=", 1)
        key = key.strip()
        local.append((key, val))
    # the include is processed first so this file's own keys override it
    for key, val in local:
        if key.lower() == "global_configuration":
            inc = _strip_comment(val).strip().strip("\"'")
            parse_mcf(os.path.join(base, inc), options)
    for key, val in local:
        if key.lower() == "global_configuration":
            continue
        _store(options, key, _strip_comment(val))
    return options


def _store(options, key, val):
    append = key.endswith("+")
    if append:
        key = key[:-1].strip()
    parts = [p for p in val.split(",")] if "," in val else [val]
    values = [infer(p) for p in parts if p.strip() != ""]
    if append:
        prev = options.get(key, [])
        if not isinstance(prev, list):
            prev = [prev]
        options[key] = prev + values
    elif len(values) == 1 and "," not in val:
        options[key] = values[0]
    else:
        options[key] = values


def apply_cli(options, argv):
    for a in argv:
        if a.startswith("--") and "=" in a:
            k, v = a[2:].split("=", 1)
            _store(options, k.strip(), v)
    return options


def get_real(options, key):
    v = options[key]
    return float(v)
