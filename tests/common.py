"""Shared helpers of the test suite: golden fixtures and option conversion."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def soda_cases() -> list[str]:
    return sorted(os.path.basename(f)[5:-4] for f in glob.glob(os.path.join(GOLDEN, "soda_*.npz")))


def load_case(name: str) -> dict:
    z = np.load(os.path.join(GOLDEN, f"soda_{name}.npz"))
    d = {k: z[k] for k in z.files}
    d["opts"] = json.loads(str(z["options"]))
    d["out"] = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return d


def sampler_options(o: dict, **extra):
    from sodabayes_amd.soda import SodaOptions
    kw = dict(method=o["method"], numit=o["numit"], burnin=o["burnin"], thin=o.get("thin", 1),
              seed=o["seed"], blocksize=o.get("blocksize", 0), nblocks=o.get("nblocks", 0),
              rblocks=bool(o.get("rblocks", False)), nmix=o.get("nmix", 2),
              mix=o.get("mix", "0.0,1.0"))
    for k in ("pi", "df_va", "scale_va", "df_ve", "scale_ve", "alpha"):
        if k in o:
            kw[k] = o[k]
    kw.update(extra)
    return SodaOptions(**kw)


def oracle_options(o: dict, nind: int, nloci: int, ngroup: int = 1):
    import orc
    mix = o.get("mix", "0.0,1.0")
    mix = [float(v) for v in mix.split(",")] if isinstance(mix, str) else list(mix)
    return orc.make_opts(o["method"], nind, nloci, nmix=o.get("nmix", 2), mix=mix, ngroup=ngroup,
                         numit=o["numit"], burnin=o["burnin"], thin=o.get("thin", 1),
                         blocksize=o.get("blocksize", 0), nblocks=o.get("nblocks", 0),
                         rblocks=bool(o.get("rblocks", False)), seed=o["seed"],
                         pizero=o.get("pi", 0.99), scale_va=o.get("scale_va", 0.01),
                         scale_ve=o.get("scale_ve", 0.01), df_va=o.get("df_va", 0.01),
                         df_ve=o.get("df_ve", 0.01), dshift_block=o.get("dshift_block", -1),
                         dshift=o.get("dshift", 0.0))


def group_ids(grp):
    """1-based group numbers by first appearance (utilities.f90:276-327)."""
    seen = {}
    out = np.empty(len(grp), dtype=np.int32)
    for j, g in enumerate(grp):
        out[j] = seen.setdefault(int(g), len(seen) + 1)
    return out, len(seen)


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)) if b.size else 0.0, 1e-300)
    return float(np.max(np.abs(a - b)) / den) if a.size else 0.0
