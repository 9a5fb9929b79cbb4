"""Extracts compact golden fixtures from the reference's committed Fortran outputs.

Reads /root/reference/tests/reference/<case>/ (the reference checkout; only this script reads it) and writes, per
case, into tests/golden/<case>/:
  input.json          the case's input, verbatim
  averages.out        fortran_averages.out, verbatim (norm, <Z>, <r>, rest norm every dtAve; printed E15.8)
  bk.npz              the BK coefficients of every analysis step (fort.21 records after its 10 header records) at
                      the coefficient indices with the largest magnitudes, plus the largest magnitude of the whole
                      file (the gate's scale) and the record length
  fort25.npz          the on-the-fly S-matrix rows (fort.25) of a subset of the energy blocks: (block, v, j, mJ,
                      order, n, m) and Re S, Im S, P and the phase as printed; the total energy of every block
  stdout.json         the analysis limits (jana, nviba, nana, nstata, inchna) and the fixed specular line
  gas.npz             (cases with debug_asyp.dat) E(v, j) and the eigenvectors of the rovibrational states
  pes.npz             (cases with pes_reference.txt) the grid shape, angular grid and a subsample of V (hartree)
  fort21_header.bin   the first 10 records of fort.21 (sizes, times, incidence, the basis labels, channel
                      energies) with their gfortran record markers, verbatim
The fixtures are data (inputs and expected outputs); no reference source is copied.

Usage: python tests/golden/make_fixtures.py
"""

from __future__ import annotations

import json
import os
import re
import shutil
import struct

import numpy as np

REFERENCE = "/root/reference/tests/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["hardwall_short", "cu111_dailight_short", "cu111_dailight_short_sym", "cu111_dailight_oblique",
         "cu111_dailight_short_lowz", "scm_cu111_srp48_small"]
BK_INDICES = 96          # coefficient indices kept per case
FORT25_BLOCKS = 5        # energy blocks kept per case (evenly spread)
PES_POINTS = 1500        # subsample of the PES fixture
# set-up only cases (the PES and the gas-phase states; no propagation output)
SETUP_CASES = ["crp_cu111_srp48_small", "scm_cu111_srp48_small_rescaling_off"]


def fortran_records(path: str) -> list[bytes]:
    data = open(path, "rb").read()
    pos, out = 0, []
    while pos < len(data):
        n = struct.unpack("<i", data[pos:pos + 4])[0]
        out.append(data[pos + 4:pos + 4 + n])
        tail = struct.unpack("<i", data[pos + 4 + n:pos + 8 + n])[0]
        assert tail == n, "malformed Fortran record"
        pos += 8 + n
    return out


def extract_bk(src: str, dst: str) -> None:
    recs = fortran_records(os.path.join(src, "fortran_fort21.bin"))[10:]
    bk = np.array([np.frombuffer(r, dtype="<c16") for r in recs])
    scale = float(np.abs(bk).max())
    order = np.argsort(-np.abs(bk).max(axis=0), kind="stable")[:BK_INDICES]
    idx = np.sort(order)
    np.savez_compressed(os.path.join(dst, "bk.npz"), records=bk[:, idx], indices=idx, scale=scale,
                        n_coefficients=bk.shape[1], n_records=bk.shape[0])


def extract_fort21_header(src: str, dst: str) -> None:
    data = open(os.path.join(src, "fortran_fort21.bin"), "rb").read()
    pos = 0
    for _ in range(10):
        n = struct.unpack("<i", data[pos:pos + 4])[0]
        pos += 8 + n
    with open(os.path.join(dst, "fort21_header.bin"), "wb") as f:
        f.write(data[:pos])


ROW = re.compile(r"^ +(-?\d+) +(-?\d+) +(\S+) +(\S+) +(\S+) +(\S+) *$")
HEADER = re.compile(r"^ v j mj dif\.order are: +(-?\d+) +(-?\d+) +(-?\d+) +(-?\d+) *$")
ENERGY = re.compile(r"^ total energy = +(\S+) *$")


def extract_fort25(src: str, dst: str) -> None:
    energies, rows, strings = [], [], []
    state = None
    for line in open(os.path.join(src, "fortran_fort25.txt")):
        line = line.rstrip("\n")
        m = ENERGY.match(line)
        if m:
            energies.append(m.group(1))
            state = None
            continue
        m = HEADER.match(line)
        if m:
            state = tuple(int(g) for g in m.groups())
            continue
        m = ROW.match(line)
        if m and state is not None and energies:
            rows.append((len(energies) - 1, *state, int(m.group(1)), int(m.group(2))))
            strings.append(m.groups()[2:])
    n_blocks = len(energies)
    keep = sorted(set(np.linspace(0, n_blocks - 1, FORT25_BLOCKS).round().astype(int).tolist()))
    sel = [i for i, r in enumerate(rows) if r[0] in keep]
    np.savez_compressed(os.path.join(dst, "fort25.npz"), rows=np.array([rows[i] for i in sel], dtype=np.int32),
                        values=np.array([strings[i] for i in sel], dtype="U24"),
                        total_energy=np.array(energies, dtype="U24"), blocks=np.array(keep, dtype=np.int32),
                        n_rows_all=len(rows))


def extract_stdout(src: str, dst: str) -> None:
    limits = {}
    fixed = []
    for line in open(os.path.join(src, "fortran_stdout.txt"), errors="replace"):
        for name in ("jana", "nviba", "nana", "nstata", "inchna"):
            m = re.search(r"(?:^|[ ,])" + name + r" = +(-?\d+)", line)
            if m:
                limits.setdefault(name, []).append(int(m.group(1)))
        if "ZAnalysis on specular channel fixed at Z =" in line:
            fixed.append(line.split()[-1])
    json.dump({"limits": limits, "fixed_line": fixed}, open(os.path.join(dst, "stdout.json"), "w"), indent=1)


def fortran_float(text: str) -> float:
    return float(text.replace("D", "E").replace("d", "e"))


def extract_gas(src: str, dst: str) -> None:
    path = os.path.join(src, "debug_asyp.dat")
    if not os.path.exists(path):
        return
    text = open(path).read()
    blocks = re.findall(r"v :\s+(\d+)\s+j:\s+(\d+)\s+E:\s+(\S+)\s+Eigenstate:\s+([^=]+)", text)
    v = np.array([int(b[0]) for b in blocks], dtype=np.int32)
    j = np.array([int(b[1]) for b in blocks], dtype=np.int32)
    e = np.array([fortran_float(b[2]) for b in blocks])
    vec = np.array([[fortran_float(x) for x in b[3].split()] for b in blocks])
    np.savez_compressed(os.path.join(dst, "gas.npz"), v=v, j=j, energy=e, vectors=vec)


def extract_pes(src: str, dst: str) -> None:
    path = os.path.join(src, "pes_reference.txt")
    if not os.path.exists(path):
        return
    lines = [ln for ln in open(path) if ln.strip() and not ln.startswith("#")]
    dims = [int(x) for x in lines[0].split()]
    sym = [int(x) for x in lines[1].split()]
    nth, nph = dims[4], dims[5]
    thetas = np.array([fortran_float(x) for x in lines[2:2 + nth]])
    phis = np.array([fortran_float(x) for x in lines[2 + nth:2 + nth + nph]])
    pts = [ln.split() for ln in lines[2 + nth + nph:]]
    idx = np.array([[int(x) - 1 for x in p[:6]] for p in pts], dtype=np.int32)
    val = np.array([fortran_float(p[6]) for p in pts])
    keep = np.linspace(0, len(pts) - 1, min(PES_POINTS, len(pts))).round().astype(int)
    np.savez_compressed(os.path.join(dst, "pes.npz"), dims=np.array(dims, dtype=np.int32),
                        symmetry=np.array(sym, dtype=np.int32), thetas=thetas, phis=phis, index=idx[keep],
                        value=val[keep])


def main() -> None:
    for case in CASES:
        src = os.path.join(REFERENCE, case)
        dst = os.path.join(HERE, case)
        os.makedirs(dst, exist_ok=True)
        shutil.copyfile(os.path.join(src, "input.json"), os.path.join(dst, "input.json"))
        shutil.copyfile(os.path.join(src, "fortran_averages.out"), os.path.join(dst, "averages.out"))
        extract_bk(src, dst)
        extract_fort21_header(src, dst)
        extract_fort25(src, dst)
        extract_stdout(src, dst)
        extract_gas(src, dst)
        extract_pes(src, dst)
        print(case, sorted(os.listdir(dst)))
    for case in SETUP_CASES:
        src = os.path.join(REFERENCE, case)
        dst = os.path.join(HERE, case)
        os.makedirs(dst, exist_ok=True)
        shutil.copyfile(os.path.join(src, "input.json"), os.path.join(dst, "input.json"))
        extract_gas(src, dst)
        extract_pes(src, dst)
        print(case, sorted(os.listdir(dst)))


if __name__ == "__main__":
    main()
