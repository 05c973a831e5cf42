"""Error metrics of a table against a reference table.

rel_err(x, y) = |x - y| / max(|y|, floor * max|y|), per element: the relative
error, except that entries below `floor` times the table's largest magnitude
are measured against that floor (a relative error of a value 1e-9 of the
table's range says nothing about the table).  floor = 1e-4 is the default;
unfloored numbers are reported alongside.  T = 0 vs 0 counts as 0.
"""
import numpy as np

FLOOR = 1e-4


def rel_err(x, y, floor=FLOOR):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    scale = np.abs(y).max() if y.size else 0.0
    den = np.maximum(np.abs(y), floor * scale)
    with np.errstate(invalid="ignore", divide="ignore"):
        e = np.where(den > 0, np.abs(x - y) / np.where(den > 0, den, 1.0), np.abs(x - y))
    return e


def summary(x, y, floor=FLOOR):
    e = rel_err(x, y, floor)
    eu = rel_err(x, y, 0.0)
    return {"max": float(e.max()) if e.size else 0.0, "mean": float(e.mean()) if e.size else 0.0,
            "max_unfloored": float(np.nanmax(eu)) if eu.size else 0.0,
            "argmax": tuple(int(i) for i in np.unravel_index(int(np.argmax(e)), e.shape)) if e.size else ()}
