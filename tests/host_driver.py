"""Helpers to run the C++ host driver executables (cabanapic_amd/host)."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cabanapic_amd", "host")
CONST_NAMES = ["qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz", "dt", "px",
               "py", "pz", "dt_eps0", "dxp", "we", "num_steps"]


def exe(deck, real):
    path = os.path.join(HOST, "build", f"cbnpic_{deck}_{real}")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", HOST], check=True)
    return path


def dump(deck, real, out, particles=True):
    flag = "--dump-initial" if particles else "--dump-constants"
    subprocess.run([exe(deck, real), flag, out], check=True, capture_output=True)
    with open(out, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        rs = int(np.frombuffer(f.read(4), np.int32)[0])
        k = np.frombuffer(f.read(16 * 8), np.float64)
        dt = np.float32 if rs == 4 else np.float64
        p = {}
        for m in ("x", "y", "z", "ux", "uy", "uz", "w"):
            p[m] = np.frombuffer(f.read(n * rs), dt).copy()
        p["cell"] = np.frombuffer(f.read(n * 4), np.int32).copy()
    return dict(zip(CONST_NAMES, k)), p
