"""Write tests/golden/hllc1d_vectors.npz: inputs and outputs of the reference's numpy-only
float64 1-D well-balanced HLLC solver (validation/hllc1d.py: solve_1d, hllc_flux), on
fully wet cases where its operator and the 2-D kernels' coincide (no dry cells, so
neither the first-order drop at dry faces nor the donor factor acts).

Run in a container that has the reference checked out (default /root/reference);
the fixture it writes is data, and the tests read only the fixture.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
sys.path.insert(0, str(REF / "validation"))

import hllc1d  # noqa: E402  (the reference module, numpy only)

OUT = Path(__file__).resolve().parent / "hllc1d_vectors.npz"


def cases():
    x = np.arange(200) + 0.5
    bump = 0.25 * np.exp(-(((x - 100.0) / 12.0) ** 2))
    yield ("flow_over_bump", dict(h0=1.2 - bump, u0=np.full(200, 0.6), z=bump, dx=1.0,
                                  t_end=20.0, manning_n=0.02))
    x = np.arange(240) + 0.5
    z = 0.002 * (240 - x) + 0.15 * np.sin(x / 9.0)
    h0 = np.where(x < 120, 1.4, 0.8) - z + 0.5
    yield ("wet_dam_break_bumpy_slope", dict(h0=h0, u0=np.zeros(240), z=z, dx=0.5,
                                             t_end=12.0, manning_n=0.015))
    x = np.arange(160) + 0.5
    z = 0.3 * np.exp(-(((x - 60.0) / 8.0) ** 2)) + 0.2 * np.exp(-(((x - 110.0) / 6.0) ** 2))
    yield ("lake_over_two_bumps", dict(h0=1.0 - z, u0=np.zeros(160), z=z, dx=2.0,
                                       t_end=30.0, manning_n=0.03))


def main() -> None:
    out = {}
    for name, c in cases():
        h, u = hllc1d.solve_1d(c["h0"], c["u0"], c["z"], c["dx"], c["t_end"], cfl=0.45,
                               manning_n=c["manning_n"])
        for k, v in c.items():
            out[f"{name}/{k}"] = np.asarray(v, dtype=np.float64)
        out[f"{name}/h"] = h
        out[f"{name}/u"] = u
    # the flux itself on a spread of face states, dry sides included
    rng = np.random.default_rng(7)
    n = 400
    hL = np.where(rng.uniform(size=n) < 0.1, 0.0, rng.uniform(0.0, 3.0, n))
    hR = np.where(rng.uniform(size=n) < 0.1, 0.0, rng.uniform(0.0, 3.0, n))
    uL, uR = rng.uniform(-4.0, 4.0, n), rng.uniform(-4.0, 4.0, n)
    vL, vR = rng.uniform(-1.0, 1.0, n), rng.uniform(-1.0, 1.0, n)
    fh, fhu, fhv = hllc1d.hllc_flux(hL, uL, vL, hR, uR, vR)
    for k, v in dict(hL=hL, uL=uL, vL=vL, hR=hR, uR=uR, vR=vR, Fh=fh, Fhu=fhu, Fhv=fhv).items():
        out[f"flux/{k}"] = np.asarray(v, dtype=np.float64)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
