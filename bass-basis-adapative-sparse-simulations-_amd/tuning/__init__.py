"""Per-size-class tuning of the kernel launches (no counterpart in the reference).

The library keeps, per size class of a call's input count n (0: n < 2^14,
1: < 2^18, 2: < 2^22, 3: larger), four knobs that change how work is laid
out.  Keys, amplitudes and kept sums never depend on them; the emission tile
size groups the input total of a dense gate output (per-tile sums, then tiles
in order), so a truncation's total -- and a kept/total ratio built from it --
may move in its last bits:

  table_factor   hash / group-table capacity = next power of two >= factor * n
                 (2: load factor <= 1/2, 4: <= 1/4)      bass_gate, bass_rdm1q
  emit_items_1q  entries per thread of a 1-qubit emission block (2, 4, 8:
                 tiles of 512 / 1024 / 2048 entries)    bass_gate (dense output)
  emit_items_2q  the same for 2-qubit gates (1, 2, 4)    bass_gate (dense output)
  sort_digits    radix digit width (8 or 10 bits)        bass_radix_sort

0 keeps the built-in choice.  autotune.py measures every value per class on
the device and writes a table; the library applies the table next to this
file (mi355x.json) when it is loaded.
"""

import json
import os

KNOBS = {"table_factor": 0, "emit_items_1q": 1, "emit_items_2q": 2, "sort_digits": 3}
VALUES = {"table_factor": (2, 4), "emit_items_1q": (2, 4, 8), "emit_items_2q": (1, 2, 4), "sort_digits": (8, 10)}
CLASS_BOUNDS = (1 << 14, 1 << 18, 1 << 22)  # upper bounds of classes 0..2; class 3 is larger
TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mi355x.json")


def size_class(n):
    for c, b in enumerate(CLASS_BOUNDS):
        if n < b:
            return c
    return len(CLASS_BOUNDS)


def validate(table):
    """Raise ValueError unless `table` maps every knob to one allowed value
    (or 0) per size class."""
    for name, allowed in VALUES.items():
        vals = table.get(name)
        if not isinstance(vals, list) or len(vals) != len(CLASS_BOUNDS) + 1:
            raise ValueError(f"tuning table: {name} needs {len(CLASS_BOUNDS) + 1} entries")
        for v in vals:
            if v != 0 and v not in allowed:
                raise ValueError(f"tuning table: {name} = {v} is not one of {allowed}")


def apply(lib, table):
    """Set the library's knobs from `table` (a dict as written by autotune.py)."""
    validate(table)
    for name, knob in KNOBS.items():
        for c, v in enumerate(table[name]):
            rc = lib.bass_tune(knob, c, int(v))
            if rc < 0:
                raise ValueError(f"bass_tune({name}, {c}, {v}) failed: {lib.bass_last_error().decode()}")


def load_table(path=TABLE_PATH):
    """The committed table, or None if there is none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)
