"""Writes the CRP fixtures under tests/golden/crp/ from the reference's CRP test data (run in the build container,
where the reference is present; the GPU box and the tests only read what this writes).

For each case: dataset/ (the data set's text files, copied unchanged: the reader is under test) and points.npz,
the inputs and outputs of the reference's Fortran CRP driver (tests/crplib/reference/<case>/*.in, *.out: IEEE
doubles as 16-digit hexadecimal, read here bit for bit):

  atom_in (n, 3) x y z / atom_out (n, 4) V, dV/dx, dV/dy, dV/dz                         (atom_0_2, bohr, hartree)
  molecule_in (n, 6) Za Xa Ya Zb Xb Yb / molecule_out (n, 7) V, dXa dYa dZa dXb dYb dZb (molecule_0_2)
  molecule13_in / molecule13_out: the same in angstrom and eV                           (molecule_1_3, when present)
  wrapper_in (n, 6) u v Z r theta phi / wrapper_out (n,) V in eV; wrapper_params (4,)   (wrapper)
  gasphase_in (n, 2) Z r / gasphase_out (n, 5) F, dF/dZ, V2D, dV2D/dZ, dV2D/dr; zso (2,) (gasphase)
  cambcord_in (n, 6) xa ya za xb yb zb / cambcord_out (n, 6) x y Z r theta phi          (cambcord)

The case cu111_srp48 uses the repository's real data set potentials/H2Cu111SRP_C3v.dir/H2Cu111SRP_C3v."""

from __future__ import annotations

import os
import shutil
import struct

import numpy as np

REF = "/root/reference"
CASES = ["cu111_srp48", "c3v_pairwise", "c6v_pairwise", "c3v_impulse_15", "c6v_impulse_5", "c4v_pairwise",
         "c4v_impulse_6", "c4v_negative_spacing"]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "crp")


def hex_double(text: str) -> float:
    return struct.unpack("<d", struct.pack("<Q", int(text, 16)))[0]


def read_in(path: str, width: int) -> np.ndarray:
    tokens = open(path).read().split()
    n = int(tokens[0])
    values = [hex_double(t) for t in tokens[1:1 + n * width]]
    return np.array(values, dtype=np.float64).reshape(n, width)


def read_out(path: str) -> list[tuple[str, float]]:
    records = []
    done = False
    for line in open(path):
        fields = line.split()
        if not fields:
            continue
        if fields[0] == "CRP_DRIVER_DONE":
            done = True
            continue
        if len(fields) >= 3 and len(fields[2]) == 16:
            records.append((fields[0], hex_double(fields[2])))
        else:
            records.append((fields[0], " ".join(fields[1:])))
    assert done, path
    return records


def grouped(records, labels: list[str], n: int) -> np.ndarray:
    out = np.zeros((n, len(labels)))
    k = 0
    for i in range(n):
        for j, label in enumerate(labels):
            assert records[k][0] == label, (records[k], label)
            out[i, j] = records[k][1]
            k += 1
    assert k == len(records)
    return out


def main() -> None:
    for case in CASES:
        src = os.path.join(REF, "tests", "crplib", "reference", case)
        dst = os.path.join(OUT, case)
        data = (os.path.join(REF, "potentials", "H2Cu111SRP_C3v.dir", "H2Cu111SRP_C3v") if case == "cu111_srp48"
                else os.path.join(src, "dataset"))
        if os.path.isdir(os.path.join(dst, "dataset")):
            shutil.rmtree(os.path.join(dst, "dataset"))
        shutil.copytree(data, os.path.join(dst, "dataset"))
        arrays = {}
        if os.path.exists(os.path.join(src, "atom_0_2.in")):
            arrays["atom_in"] = read_in(os.path.join(src, "atom_0_2.in"), 3)
            arrays["atom_out"] = grouped(read_out(os.path.join(src, "atom_0_2.out")), ["V", "F", "F", "F"],
                                         len(arrays["atom_in"]))
        for name, key in (("molecule_0_2", "molecule"), ("molecule_1_3", "molecule13")):
            if os.path.exists(os.path.join(src, name + ".out")):
                arrays[key + "_in"] = read_in(os.path.join(src, name + ".in"), 6)
                arrays[key + "_out"] = grouped(read_out(os.path.join(src, name + ".out")), ["V"] + ["F"] * 6,
                                               len(arrays[key + "_in"]))
        recs = read_out(os.path.join(src, "wrapper.out"))
        assert recs[0][0] == "point_group"
        arrays["point_group"] = np.array(recs[0][1])
        arrays["wrapper_params"] = np.array([v for _, v in recs[1:5]])
        arrays["wrapper_in"] = read_in(os.path.join(src, "wrapper.in"), 6)
        arrays["wrapper_out"] = grouped(recs[5:], ["V"], len(arrays["wrapper_in"]))[:, 0]
        if os.path.exists(os.path.join(src, "gasphase.out")):
            recs = read_out(os.path.join(src, "gasphase.out"))
            arrays["zso"] = np.array([v for _, v in recs[:2]])
            arrays["gasphase_in"] = read_in(os.path.join(src, "gasphase.in"), 2)
            arrays["gasphase_out"] = grouped(recs[2:], ["G"] * 5, len(arrays["gasphase_in"]))
        if os.path.exists(os.path.join(src, "cambcord.out")):
            arrays["cambcord_in"] = read_in(os.path.join(src, "cambcord.in"), 6)
            arrays["cambcord_out"] = grouped(read_out(os.path.join(src, "cambcord.out")), ["C"] * 6,
                                             len(arrays["cambcord_in"]))
        np.savez_compressed(os.path.join(dst, "points.npz"), **arrays)
        print(case, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
