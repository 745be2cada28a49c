"""Per-kernel timing from the library's HIP-event profiler (rama_profile_*).

Events are recorded on the stream the solver launches on, so durations are
those of the kernels themselves; bytes are each kernel's algorithmic HBM
traffic (what an ideal implementation must move), giving achieved GB/s.
"""
import ctypes

from . import _capi


def enable(on=True):
    L = _capi.lib()
    if on:
        L.rama_profile_reset()
    L.rama_profile_enable(1 if on else 0)


def only(name=None):
    """Time only launches named `name` (None: all)."""
    _capi.lib().rama_profile_only(name.encode() if name else None)


def read():
    L = _capi.lib()
    out = []
    for k in range(L.rama_profile_count()):
        name = ctypes.c_char_p()
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        by = ctypes.c_double()
        if L.rama_profile_entry(k, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)) == 0:
            d = {"name": name.value.decode(), "launches": n.value, "ms": ms.value, "bytes": by.value}
            units, entries = ctypes.c_double(), ctypes.c_double()
            if L.rama_profile_work(k, ctypes.byref(units), ctypes.byref(entries)) == 0 and units.value > 0:
                d["units"], d["entries"] = units.value, entries.value
            out.append(d)
    out.sort(key=lambda d: -d["ms"])
    return out


def top(stats, k=8):
    """The k slowest kernels: achieved GB/s of their algorithmic bytes; the
    cycle-search counts add expansions/s and neighbour-list entries/s (their
    GB/s is the merge-walk model: 4 bytes per list entry and 8 bytes of list
    bounds per expansion, served mostly from cache)."""
    out = []
    for d in stats[:k]:
        e = {"name": d["name"], "launches": d["launches"], "ms": round(d["ms"], 3),
             "GBps": round(d["bytes"] / (d["ms"] * 1e6), 1) if d["bytes"] > 0 and d["ms"] > 0 else None}
        if d.get("units") and d["ms"] > 0:
            e["expansions_per_s"] = d["units"] / (d["ms"] * 1e-3)
            e["entries_per_s"] = d["entries"] / (d["ms"] * 1e-3)
        out.append(e)
    return out


def dominant(stats):
    """Name of the kernel with a byte model that takes the most time."""
    cands = [d for d in stats if d["bytes"] > 0 and d["ms"] > 0 and not d.get("units")]
    return max(cands, key=lambda x: x["ms"])["name"] if cands else None


def roofline(stats, peak_gbs, name=None):
    """Roofline entry of the dominant kernel that has a byte model (or `name`)."""
    cands = [d for d in stats if d["bytes"] > 0 and d["ms"] > 0 and not d.get("units")
             and (name is None or d["name"] == name)]
    if not cands:
        return None
    d = max(cands, key=lambda x: x["ms"])
    per_launch_bytes = d["bytes"] / d["launches"]
    per_launch_s = d["ms"] * 1e-3 / d["launches"]
    achieved = per_launch_bytes / per_launch_s / 1e9
    return {"kernel": d["name"], "bound": "hbm", "achieved": round(achieved, 1), "peak": peak_gbs, "unit": "GB/s",
            "frac": round(achieved / peak_gbs, 4), "traffic": None, "launches": d["launches"],
            "avg_launch_us": round(per_launch_s * 1e6, 2), "algorithmic_bytes_per_launch": per_launch_bytes}


def traffic_table(path):
    """Per-kernel HBM traffic (bytes per dispatch) from a committed rocprofv3
    counter summary: the markdown table whose rows read
    | profiler name | kernel | dispatches | fetch | write | traffic |."""
    import os
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[2].isdigit():
            out[cells[0]] = float(cells[5])
    return out
