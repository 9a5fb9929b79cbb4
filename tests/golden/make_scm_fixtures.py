"""Writes the SCM fixtures under tests/golden/scm/ from the reference's SCM test data (run in the build container,
where the reference is present; the tests only read what this writes).

Data files, copied unchanged (their readers are under test): example/params.dat and example/examplePosfile.dat
(the reference's potentials/H2Cu111SRP_C3v.dir), synthetic/params.in, slab_v2.in, slab_v3.in
(tests/scm/reference/synthetic). For each of the reference's four Fortran driver runs
(tests/scm/scm_reference_runs.hpp:29-38) <run>.npz holds, bit for bit from the driver output:
  points (n, 6) u v Z r theta phi; out (n, 5) x, y, delta_v (hartree), v_crp (eV), v_total (eV);
  scalars: rescaling (0/1), slab_number, cutoff, scale, cos_skew, sin_skew, lattice_x, atoms (the atoms in use)."""

from __future__ import annotations

import os
import shutil
import struct

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scm")
RUNS = [("example", "rescaling_on"), ("example", "rescaling_off"),
        ("synthetic", "v3_slab3_rescaling_on"), ("synthetic", "v2_slab2_rescaling_off")]


def hex_double(text: str) -> float:
    return struct.unpack("<d", struct.pack("<Q", int(text, 16)))[0]


def main() -> None:
    os.makedirs(os.path.join(OUT, "example"), exist_ok=True)
    os.makedirs(os.path.join(OUT, "synthetic"), exist_ok=True)
    srp = os.path.join(REF, "potentials", "H2Cu111SRP_C3v.dir")
    for name in ("params.dat", "examplePosfile.dat"):
        shutil.copyfile(os.path.join(srp, name), os.path.join(OUT, "example", name))
    for name in ("params.in", "slab_v2.in", "slab_v3.in"):
        shutil.copyfile(os.path.join(REF, "tests", "scm", "reference", "synthetic", name),
                        os.path.join(OUT, "synthetic", name))
    for directory, run in RUNS:
        src = os.path.join(REF, "tests", "scm", "reference", directory)
        tokens = open(os.path.join(src, "points.in")).read().split()
        n = int(tokens[0])
        points = np.array([hex_double(t) for t in tokens[1:1 + 6 * n]]).reshape(n, 6)
        values: dict[str, list[float]] = {}
        text: dict[str, str] = {}
        atoms = 0
        for line in open(os.path.join(src, run + ".out")):
            f = line.split()
            if not f or f[0] == "SCM_DRIVER_DONE":
                continue
            if len(f) >= 3 and len(f[2]) == 16 and all(c in "0123456789ABCDEFabcdef" for c in f[2]):
                values.setdefault(f[0], []).append(hex_double(f[2]))
            elif f[0] == "atom_layer":
                atoms += 1
            else:
                text[f[0]] = " ".join(f[1:])
        out = np.column_stack([values[k] for k in ("x", "y", "delta_v", "v_crp", "v_total")])
        assert out.shape == (n, 5), (run, out.shape)
        np.savez_compressed(os.path.join(OUT, directory, run + ".npz"), points=points, out=out,
                            rescaling=int(text["rescaling"] == "T"), slab_number=int(text["slab_number"]),
                            cutoff=values["cutoff"][0], scale=values["scale"][0], cos_skew=values["cos_skew"][0],
                            sin_skew=values["sin_skew"][0], lattice_x=values["lattice_x"][0], atoms=atoms)
        print(directory, run, n, "points", atoms, "atoms")


if __name__ == "__main__":
    main()
