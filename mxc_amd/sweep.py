"""The serial experiment sweep of the reference (scripts/run.sh:12-18): lorasp
at N = 2048 .. 32768 with theta 2, leaf 256, eps 1e-10, rank 100 and sampling
2000, one run after another, every report appended to serial.log and every
result row to serial.csv (write_to_csv's 25 columns plus the GMRES residual
history, test_funcs.hpp:94-111). The reference turns its logs into a CSV with
a helper (scripts/run.sh:31-34); lorasp writes the row itself here. The
strong-scaling half of run.sh (mpirun -n 1..16) needs more than one GPU and
is not part of this build.

    python -m mxc_amd.sweep --out DIR [--sizes 2048,4096] [-- extra lorasp options]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys

LORASP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "lorasp")
SERIAL_SIZES = (2048, 4096, 8192, 16384, 32768)
SERIAL_ARGS = ("2", "256", "1.e-10", "100", "2000")  # theta leaf eps rank sampling


def serial(out_dir: str, sizes=SERIAL_SIZES, extra=(), timeout: float = 1200.0):
    """Runs the serial sweep into out_dir (created; an existing serial.log
    and serial.csv are replaced, as run.sh starts from an empty log folder).
    Returns the paths of the log and the CSV."""
    if not os.path.exists(LORASP):
        raise FileNotFoundError(f"{LORASP} is not built (run __graft_entry__.build())")
    os.makedirs(out_dir, exist_ok=True)
    log, csv = os.path.join(out_dir, "serial.log"), os.path.join(out_dir, "serial.csv")
    for p in (log, csv):
        if os.path.exists(p):
            os.remove(p)
    for n in sizes:
        cmd = [LORASP, str(n), *SERIAL_ARGS, "--csv", csv, *extra]
        print("+ " + " ".join(cmd), flush=True)
        with open(log, "a") as f:
            r = subprocess.run(cmd, stdout=f, stderr=subprocess.PIPE, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"lorasp {n} failed ({r.returncode}): {r.stderr.strip()}")
    return log, csv


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="folder of serial.log and serial.csv")
    ap.add_argument("--sizes", default=",".join(map(str, SERIAL_SIZES)))
    ap.add_argument("extra", nargs="*", help="further lorasp options (after --)")
    a = ap.parse_args(argv)
    log, csv = serial(a.out, [int(s) for s in a.sizes.split(",") if s], a.extra)
    print(f"sweep done: {log}, {csv}")


if __name__ == "__main__":
    sys.exit(main())
