"""Parameter sweep from the command line on one resident batch.

The reference sweeps pump power and waist by running its executable once per
point from a shell loop (src/cuOPO3D.sh:56-86: `./cuOPO3D $N $WAIST`, N in
mW, output moved to sPPLT_N_<N>/ under waist_<W>/).  Here every point is a
solve of one batch (opo_batch_*): all points advance together, one launch
per split-step stage per size class, and each point's outputs land in the
same folder layout:

    OUT/waist_<W>/sPPLT_N_<N>/power_per_trip.dat      trip, Pp, Ps, Pi
    OUT/waist_<W>/sPPLT_N_<N>/OutputSignal..._r.dat   x-y slice NT/2 (and _i)
    OUT/waist_<W>/sPPLT_N_<N>/As_temporal_..._r.dat   time trace, centre pixel
    OUT/waist_<W>/sPPLT_N_<N>/{pump,signal,idler}.npy  whole fields (--npy)

Crystal and cavity follow src/cuOPO3D.cu:19-46 (MgO:sPPLT, 29 mm, 27 C,
Lambda = 6.97 um, Lcav = 4 Lcr, Rs = Ri = 0.98, GDD compensation, focused
CW pump focused at Lcr/2, signal and idler resonant, Ai = As noise).

usage: python -m cuopo3d_amd.sweep --powers 500:600:50 --waists 31 \
           [--grid 64 256] [--trips 200] [--out DIR] [--budget-gb 128] [--npy]
Ranges are START:STOP:STEP inclusive, like the awk loops of cuOPO3D.sh.
"""
import argparse
import math
import os
import sys

import numpy as np

from . import Batch, configs
from .capi import Grid, OPO_FP32


def parse_range(text):
    """'a:b:s' -> [a, a+s, ..., <= b]; 'a' -> [a]; 'a,b,c' -> list."""
    if "," in text:
        return [float(v) for v in text.split(",")]
    parts = [float(v) for v in text.split(":")]
    if len(parts) == 1:
        return parts
    a, b = parts[0], parts[1]
    s = parts[2] if len(parts) > 2 else 1.0
    if s <= 0:
        raise ValueError("step must be positive")
    n = int(math.floor((b - a) / s + 1e-9)) + 1
    return [a + k * s for k in range(max(n, 0))]


def reference_physical(n_round_trips, precision=OPO_FP32):
    """Crystal and cavity of src/cuOPO3D.cu:19-46."""
    ph = configs.physical(lcr=29.0e3, nz=100, lx=1.0e3, rs=0.98, ri=0.98, resonant_s=1, resonant_i=1,
                          gdd_on=1, gdd_gamma=1.0, fs_on=1, n_round_trips=n_round_trips,
                          precision=precision, dq=0.0)
    ph.crystal = configs.CRYSTAL_MGOSPPLT
    return ph


def point(grid, ph, power_mw, waist_um):
    p = configs.make_params(grid, ph)
    api, as_, ai = configs.make_fields(grid, p, ph, mode=configs.PUMP_FOCUSED_CW, power=power_mw * 1e-3,
                                       waist=waist_um, seed=1)
    return p, (api, as_, ai)


def _fmt(v):
    return f"{v:g}"


def write_dat(path_stem, values, rows):
    """Reference .dat text: 20 significant digits, tab after each value
    (Files.cuh:73-94); rows=True writes a 2-D slice, else one value per line."""
    for part, ext in ((np.real, "_r.dat"), (np.imag, "_i.dat")):
        data = part(values)
        with open(path_stem + ext, "w") as fh:
            if rows:
                for row in data:
                    fh.write("".join(f"{v:.20g}\t" for v in row) + "\n")
            else:
                fh.write("".join(f"{v:.20g}\n" for v in data))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--powers", default="500", help="pump powers in mW, START:STOP:STEP")
    ap.add_argument("--waists", default="31", help="waists in um, START:STOP:STEP")
    ap.add_argument("--grid", type=int, nargs=2, default=(64, 256), metavar=("NXY", "NT"))
    ap.add_argument("--trips", type=int, default=200)
    ap.add_argument("--out", default=".")
    ap.add_argument("--budget-gb", type=float, default=128.0)
    ap.add_argument("--npy", action="store_true", help="also save the three whole fields as .npy")
    args = ap.parse_args(argv)

    grid = Grid(args.grid[0], args.grid[0], args.grid[1])
    ph = reference_physical(args.trips)
    b = Batch(args.budget_gb * (1 << 30))
    jobs = {}
    for w in parse_range(args.waists):
        for pw in parse_range(args.powers):
            p, f = point(grid, ph, pw, w)
            t = b.submit(grid, p, *f, args.trips)
            jobs[t] = (pw, w)
    print(f"sweep: {len(jobs)} points on {grid.nx}x{grid.ny}x{grid.nt}, {args.trips} round trips",
          flush=True)
    done = {}
    while len(done) < len(jobs):
        b.step()
        for t, (pw, w) in jobs.items():
            if t in done:
                continue
            r = b.poll(t)
            if r is None:
                continue
            power, (ap_, as_, ai_) = r
            folder = os.path.join(args.out, f"waist_{_fmt(w)}", f"sPPLT_N_{_fmt(pw)}")
            os.makedirs(folder, exist_ok=True)
            with open(os.path.join(folder, "power_per_trip.dat"), "w") as fh:
                for k, row in enumerate(power):
                    fh.write(f"{k}\t{row[0]:.17g}\t{row[1]:.17g}\t{row[2]:.17g}\n")
            tag = f"_PumpPower{_fmt(pw)}_W"
            write_dat(os.path.join(folder, f"OutputSignalPower{tag}_slice{grid.nt // 2}"),
                      as_[grid.nt // 2], rows=True)
            c = grid.nx // 2
            write_dat(os.path.join(folder, f"As_temporal_x_{c}_y_{c}"), as_[:, c, c], rows=False)
            if args.npy:   # shape (NT, NY, NX), complex64, the solver's [t][y][x] layout
                for name, fld in (("pump", ap_), ("signal", as_), ("idler", ai_)):
                    np.save(os.path.join(folder, f"{name}.npy"), fld)
            done[t] = power[-1, 1]
            print(f"  P = {_fmt(pw)} mW, w0 = {_fmt(w)} um: signal power {power[-1, 1]:.6e} W -> {folder}",
                  flush=True)
    b.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
