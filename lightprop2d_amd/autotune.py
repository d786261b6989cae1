"""Autotuner: picks the row-pass, column-pass and chained-plane row-pass schedule
(elements per thread E, rows per workgroup, columns per strip W) for each grid size
and precision by timing every registered variant on the device, then installs the
winners with ``lp2d_set_variant``.

Results are cached as JSON keyed by "<gfx arch>:<dtype>:<n>".  A table measured on
MI355X ships with the package (tuned_gfx950.json) and is applied on import of
``lightprop2d_amd.propagation``; ``tune()`` re-measures and updates the user cache
(``$LP2D_TUNE_CACHE`` or ~/.cache/lightprop2d_amd/tune.json).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

from . import _lib as L

_PKG = os.path.dirname(os.path.abspath(__file__))
SHIPPED = os.path.join(_PKG, "tuned_gfx950.json")
_DT = {"c64": L.LP2D_C64, "c128": L.LP2D_C128}


def cache_path():
    return os.environ.get("LP2D_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "lightprop2d_amd",
                                       "tune.json"))


def arch() -> str:
    import torch
    try:
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:  # pragma: no cover
        return "unknown"


def variant_names(n: int, dtype: str = "c64"):
    lib = L.lib()
    buf = C.create_string_buffer(128)
    out = {}
    for kind, label in ((0, "rows"), (1, "cols"), (3, "chain")):
        names = []
        for v in range(lib.lp2d_variant_count(n, _DT[dtype], kind)):
            lib.lp2d_variant_name(n, _DT[dtype], kind, v, buf, 128)
            names.append(buf.value.decode())
        out[label] = names
    return out


def _collect(lib):
    arr = (L.lp2d_trace_stat * 64)()
    cnt = C.c_int()
    L.check(lib.lp2d_trace_collect(arr, 64, C.byref(cnt)), "trace_collect")
    return {arr[i].name.decode(): arr[i].total_ms / max(1, arr[i].launches)
            for i in range(cnt.value)}


def _time_variant(lib, step, reps):
    import torch
    step()
    torch.cuda.synchronize()
    _collect(lib)
    L.check(lib.lp2d_trace_enable(1))
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    L.check(lib.lp2d_trace_enable(0))
    return _collect(lib)


def tune(n: int, dtype: str = "c64", batch: int | None = None, reps: int = 5,
         save: bool = True, verbose: bool = False) -> dict:
    """Measure every row/column variant at size n and install the fastest."""
    import torch
    from . import propagation as P
    lib = L.lib()
    dt = _DT[dtype]
    tdt = torch.complex64 if dtype == "c64" else torch.complex128
    if batch is None:
        batch = max(1, min(64, (256 << 20) // (n * n * (8 if dtype == "c64" else 16))))
    u = torch.randn((batch, n, n), dtype=tdt, device="cuda")
    out = torch.empty_like(u)

    def step():
        P.propagate(u, 0.01, 400e-4, 632e-7, out=out)

    result = {"n": n, "dtype": dtype, "batch": batch, "rows": [], "cols": []}
    if P._small_path(n, dtype):
        result["note"] = "whole field in LDS: one kernel, no variants"
        return result
    nr = lib.lp2d_variant_count(n, dt, 0)
    nc = lib.lp2d_variant_count(n, dt, 1)
    names = variant_names(n, dtype)
    L.check(lib.lp2d_set_variant(n, dt, 1, 0))
    best_r, best_rt = 0, float("inf")
    for v in range(nr):
        L.check(lib.lp2d_set_variant(n, dt, 0, v))
        t = _time_variant(lib, step, reps)
        rt = sum(ms for k, ms in t.items() if k.startswith("rows_"))
        result["rows"].append({"variant": v, "name": names["rows"][v], "ms": rt})
        if rt < best_rt:
            best_r, best_rt = v, rt
    L.check(lib.lp2d_set_variant(n, dt, 0, best_r))
    best_c, best_ct = 0, float("inf")
    for v in range(nc):
        L.check(lib.lp2d_set_variant(n, dt, 1, v))
        t = _time_variant(lib, step, reps)
        ct = sum(ms for k, ms in t.items() if k.startswith("cols_"))
        result["cols"].append({"variant": v, "name": names["cols"][v], "ms": ct})
        if ct < best_ct:
            best_c, best_ct = v, ct
    L.check(lib.lp2d_set_variant(n, dt, 1, best_c))
    result["best"] = {"rows": best_r, "cols": best_c}
    entry = {"rows": best_r, "cols": best_c, "rows_name": names["rows"][best_r],
             "cols_name": names["cols"][best_c], "rows_ms": best_rt, "cols_ms": best_ct,
             "batch": batch, "time": time.time()}
    # chained-plane row pass of lp2d_propagate_stack, timed on a 3-screen stack
    nch = lib.lp2d_variant_count(n, dt, 3)
    if nch > 0:
        scr = torch.rand((3, n, n), device="cuda", dtype=torch.float32)

        def stack_step():
            P.propagate_stack(u, scr, 0.01, 400e-4, 632e-7, out=out)

        result["chain"] = []
        best_h, best_ht = 0, float("inf")
        for v in range(nch):
            L.check(lib.lp2d_set_variant(n, dt, 3, v))
            t = _time_variant(lib, stack_step, reps)
            ht = sum(ms for k, ms in t.items() if k.startswith("rows_chain"))
            result["chain"].append({"variant": v, "name": names["chain"][v], "ms": ht})
            if ht < best_ht:
                best_h, best_ht = v, ht
        L.check(lib.lp2d_set_variant(n, dt, 3, best_h))
        result["best"]["chain"] = best_h
        entry.update({"chain": best_h, "chain_name": names["chain"][best_h],
                      "chain_ms": best_ht})
    if verbose:
        print(json.dumps(result))
    if save:
        _save({f"{arch()}:{dtype}:{n}": entry})
    return result


def _save(entries: dict):
    p = cache_path()
    os.makedirs(os.path.dirname(p), exist_ok=True)
    data = {}
    if os.path.exists(p):
        try:
            data = json.load(open(p))
        except (OSError, ValueError):
            data = {}
    data.update(entries)
    with open(p + ".tmp", "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(p + ".tmp", p)


def apply(table: dict, gfx: str | None = None) -> int:
    """Install the variants of a {"<arch>:<dtype>:<n>": {"rows": r, "cols": c}} table
    for the given arch; returns how many sizes were set."""
    lib = L.lib()
    k = 0
    for key, v in table.items():
        a, dtype, n = key.split(":")
        if gfx is not None and a != gfx:
            continue
        n = int(n)
        names = variant_names(n, dtype)
        for kind, name in ((0, "rows"), (1, "cols"), (3, "chain")):
            if name not in v and name + "_name" not in v:
                continue  # (tables without a chained-pass entry keep its default)
            want = v.get(name + "_name")
            if want is not None:  # match by schedule, not by index
                if want in names[name]:
                    L.check(lib.lp2d_set_variant(n, _DT[dtype], kind, names[name].index(want)))
            elif lib.lp2d_variant_count(n, _DT[dtype], kind) > v[name]:
                L.check(lib.lp2d_set_variant(n, _DT[dtype], kind, int(v[name])))
        k += 1
    return k


def apply_defaults(gfx: str = "gfx950") -> int:
    """Apply the shipped table, then the user cache (which wins)."""
    k = 0
    for p in (SHIPPED, cache_path()):
        if os.path.exists(p):
            try:
                k += apply(json.load(open(p)), gfx)
            except (OSError, ValueError, KeyError):
                pass
    return k
