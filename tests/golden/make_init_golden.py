"""Golden vectors of the reference's scenario setup: runs the reference's
own initialize (in_out/initialize.f90 particle_placing from the domain map,
boundaries.f90 boundary_placing_1 from the boundary map, the VP time step)
through oracle/_ref/simp_ref_init (built from the reference sources by
oracle/build_ref.sh; runs in the build container only) and writes
tests/golden/reference_default_init.npz: the two maps and their geometry
(the generator's inputs) and the initial state they produce."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "simp_ref_init")


def main():
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: run oracle/build_ref.sh first")
    with tempfile.TemporaryDirectory() as tmp:
        os.makedirs(os.path.join(tmp, "out"))
        fout = os.path.join(tmp, "init.bin")
        subprocess.run([EXE, fout, "/out/"], cwd=tmp, env=dict(os.environ, PWD=tmp),
                       check=True, stdout=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    n, nb, mx, my, bx, by = take("<i4", 6)
    mp = take("<i4", mx * my).reshape(my, mx).T        # Fortran (x, y) order
    bmp = take("<i4", bx * by).reshape(by, bx).T
    geo = take("<f8", 11)
    dt, max_hsml, time = take("<f8", 3)
    x = take("<f8", 2 * n).reshape(n, 2)
    vx = take("<f8", 2 * n).reshape(n, 2)
    mass, rho, hsml, A, h = (take("<f8", n) for _ in range(5))
    xb = take("<f8", 2 * nb).reshape(nb, 2)
    massb = take("<f8", nb)
    assert off == len(raw)
    np.savez_compressed(os.path.join(HERE, "reference_default_init.npz"),
                        map=mp, bmap=bmp, geometry=geo, dt=dt, max_hsml=max_hsml, time=time,
                        x=x, vx=vx, mass=mass, rho=rho, hsml=hsml, A=A, h=h, xb=xb,
                        massb=massb)
    print("written", n, "particles,", nb, "boundary particles")


if __name__ == "__main__":
    main()
