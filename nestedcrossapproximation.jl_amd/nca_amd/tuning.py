"""Persisted autotuning: a JSON table of launch choices by device and
operator shape, so that a process applies the choices measured earlier
(nca_h2_tuning_apply) instead of timing every candidate again
(nca_h2_autotune).

The key of an operator holds the device architecture, the precision and the
shape of the operator (matrix size, coupling blocks and entries, near pairs
and entries, coefficients per side): operators with the same key run the same
kernels over the same launch plans. The table shipped with the package
(``tuning_mi355x.json``) was measured on MI355X for the benchmark workloads.
"""
import ctypes
import json
import os

from . import _abi
from ._native import hip_lib
from .runtime import call

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")

# the launch choices of nca_h2_tuning (timing fields are not part of an entry)
CHOICES = ("near_k3", "leaf_order", "near_xcd_order", "near_stagger", "couple_k3", "couple_tile_order",
           "couple_xcd_order")


def device_name(ctx):
    buf = ctypes.create_string_buffer(256)
    call(hip_lib().nca_ctx_device_name(ctx.handle, buf, len(buf)))
    return buf.value.decode()


def operator_key(core, precision):
    s = core.stats()
    return "|".join([
        device_name(core.ctx).split(":")[0], "mixed" if precision == _abi.NCA_PREC_MIXED else "fp64",
        f"M{core.dim[0]}", f"N{core.dim[1]}", f"cb{s['coupling_blocks']}", f"ce{s['coupling_entries']}",
        f"np{s['near_pairs']}", f"ne{s['near_entries']}", f"ct{s['coeffs_test']}", f"cs{s['coeffs_trial']}"])


class TuningTable:
    """Launch choices by operator key, read from and written to a JSON file."""

    def __init__(self, path=DEFAULT_TABLE):
        self.path = path
        self.entries = {}
        if path and os.path.exists(path):
            with open(path) as fh:
                self.entries = json.load(fh)

    def get(self, key):
        return self.entries.get(key)

    def put(self, key, tuning, workload=None):
        e = {n: int(tuning[n]) for n in CHOICES}
        e["near_ms"] = round(float(tuning["near_ms"]), 4)
        e["couple_ms"] = round(float(tuning["couple_ms"]), 4)
        if workload:
            e["workload"] = workload
        self.entries[key] = e

    def save(self, path=None):
        path = path or self.path
        with open(path, "w") as fh:
            json.dump(self.entries, fh, indent=1, sort_keys=True)
            fh.write("\n")


def apply(core, entry):
    """Set the operator's launch choices for its current precision."""
    t = _abi.H2Tuning()
    for n in CHOICES:
        setattr(t, n, int(entry[n]))
    call(hip_lib().nca_h2_tuning_apply(core.handle, ctypes.byref(t)))


def tune(core, precision, table=None, reps=3, workload=None):
    """The operator's launch choices for ``precision`` (which it must be set
    to): from ``table`` when it holds the operator's key, else measured by
    nca_h2_autotune and added to the table. Returns the choices with
    ``source`` "table" or "measured"."""
    key = operator_key(core, precision)
    if table is not None:
        e = table.get(key)
        if e is not None:
            apply(core, e)
            return dict(e, source="table", key=key)
    r = core.autotune(reps=reps)
    if table is not None:
        table.put(key, r, workload)
    return dict(r, source="measured", key=key)
