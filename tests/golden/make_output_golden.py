"""Golden CSV snapshots written by the reference's own output routine
(in_out/output.f90) through oracle/_ref/simp_ref_output (built from the
reference sources by oracle/build_ref.sh; runs in the build container
only).  Writes tests/golden/output_csv/<case>/{input.npz, particle_*.csv}.

Cases:
  edge     -- 12 particles of format edge values (signed zeros, ties at the
              ninth digit, a value rounding up to the next decade, a
              denormal, 1e300, NaN, +-Inf) and random ones; two snapshots
  box500   -- the first 500 particles of the reference's state after 20
              predictor-corrector steps (box2500_pc20.npz), one snapshot
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "simp_ref_output")
FIELDS = ("mass", "p", "hsml", "A", "h")


def write_case(name, n, nb, snaps):
    """snaps: list of dicts with time, x (n,2), vx (n,2), mass, p, hsml, A, h,
    eps (n,3), sigma (n,3)."""
    out = os.path.join(HERE, "output_csv", name)
    os.makedirs(out, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        fin = os.path.join(tmp, "in.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<3i", n, nb, len(snaps)))
            for s in snaps:
                f.write(struct.pack("<d", s["time"]))
                arrs = [s["x"], s["vx"]] + [s[k] for k in FIELDS] + \
                       [s["eps"][:, c] for c in range(3)] + [s["sigma"][:, c] for c in range(3)]
                for a in arrs:
                    f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        os.makedirs(os.path.join(tmp, "out"))
        env = dict(os.environ, PWD=tmp)
        subprocess.run([EXE, fin, "/out/"], cwd=tmp, env=env, check=True)
        for fn in ("particle_momentum.csv", "particle_state.csv", "particle_other.csv"):
            with open(os.path.join(tmp, "out", fn), "rb") as f:
                data = f.read()
            with open(os.path.join(out, fn), "wb") as f:
                f.write(data)
    flat = {"n": n, "nb": nb, "snapshots": len(snaps)}
    for i, s in enumerate(snaps):
        for k, v in s.items():
            flat[f"s{i}_{k}"] = np.asarray(v)
    np.savez_compressed(os.path.join(out, "input.npz"), **flat)


def edge_case():
    rng = np.random.default_rng(20)
    n = 12
    special = np.array([0.0, -0.0, 1.0, -1.0, 123456785.0, 123456775.0, -123456785.0,
                        9.9999999995e5, 5e-324, 1e300, np.nan, np.inf])
    snaps = []
    for t, flip in ((7200.0, 1.0), (9000.5, -1.0)):
        s = {"time": t,
             "x": np.stack([special * 1000 * flip, rng.normal(0, 3e5, n)], 1),
             "vx": np.stack([special / 100, rng.normal(0, 0.2, n)], 1),
             "mass": np.abs(rng.normal(0, 1e9, n)),
             "p": -np.abs(special) * flip,
             "hsml": special[::-1].copy(),
             "A": rng.uniform(0, 1, n),
             "h": rng.lognormal(0, 1, n),
             "eps": rng.normal(0, 1e-7, (n, 3)),
             "sigma": rng.normal(0, 1e4, (n, 3))}
        snaps.append(s)
    write_case("edge", n, 3, snaps)


def box_case():
    g = np.load(os.path.join(HERE, "box2500_pc20.npz"))
    n = 500
    s = {"time": float(g["out_time"]), "x": g["out_x"][:n], "vx": g["out_vx"][:n],
         "mass": g["in_mass"][:n], "p": g["out_p"][:n], "hsml": g["out_hsml"][:n],
         "A": g["out_A"][:n], "h": g["out_h"][:n], "eps": g["out_eps"][:n],
         "sigma": g["out_sigma"][:n]}
    write_case("box500", n, int(g["in_xb"].shape[0]), [s])


if __name__ == "__main__":
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: run oracle/build_ref.sh first")
    edge_case()
    box_case()
    print("written", os.path.join(HERE, "output_csv"))
