"""Loaders for the committed golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mwu_cases():
    z = np.load(os.path.join(GOLDEN, "mwu_cases.npz"))
    out = []
    for name in z["__names__"]:
        name = str(name)
        d = {k.split("__", 1)[1]: z[k] for k in z.files if k.startswith(name + "__")}
        for key in ("n_rows", "n_targets"):
            d[key] = int(d[key])
        for key in ("tie_correction", "use_continuity"):
            d[key] = bool(d[key])
        for key in ("zero_handling", "alternative", "method"):
            d[key] = str(d[key])
        d["name"] = name
        out.append(d)
    return out


def group_mean_cases():
    z = np.load(os.path.join(GOLDEN, "group_mean_cases.npz"))
    out = []
    for k in range(4):
        p = f"gm{k}__"
        out.append(dict(data=z[p + "data"], indices=z[p + "indices"], indptr=z[p + "indptr"],
                        group_id=z[p + "group_id"], n_groups=int(z[p + "n_groups"]),
                        include_zeros=bool(z[p + "include_zeros"]), expected=z[p + "expected"]))
    return out


def small_config_expected():
    z = np.load(os.path.join(GOLDEN, "small_config_expected.npz"))
    return {k: z[k] for k in z.files}


def rel_close(a, b, rtol=1e-12, atol=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(a), np.abs(b)) + atol)
