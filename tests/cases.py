"""Input files of the test and benchmark workloads (the reference's JSON keys).

GOLDEN_DIR holds the inputs of the reference's Fortran-gated cases, copied as data together with the extracted
Fortran outputs (tests/golden/make_fixtures.py). The dictionaries below are this repository's own workloads:
  SMOKE_DAILIGHT  a small Dai-Light H2/Cu(111) grid with full angular coupling (seconds on the CPU oracle);
  DZ_ANGULAR      the same grid with the DZ workload's angular basis (JMax 24, MJMax 23, C6v);
  HARDWALL_6D     the HardWallPES workload: NZ 128 on Z in [0.5, 20], NR 32 on r in [0.6, 4.5], 8 x 8 (x, y),
                  JMax 10, 2000 steps of 2 a.u.;
  DZ_6D           the Dai-Light H2/Cu(111) LEPS workload: NZ 96 on Z in [-1, 12], NR 48 on r in [0.8, 6],
                  16 x 16 (x, y), JMax 24 with MJMax 23 (13 theta x 48 phi), switched to the gas phase between
                  Z = 6 and 8, C6v symmetry-reduced potential, specular grid of 512 Z points.
"""

from __future__ import annotations

import json
import os
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_DIR = os.path.join(HERE, "golden")

_COMMON = {
    "NewCalc": True,
    "ZDoF": True, "RDoF": True, "XDoF": True, "YDoF": True, "ThetaDoF": True, "PhiDoF": True,
    "Mass1": 1.0078, "Mass2": 1.0078, "HomoNuc": True,
    "ProjOp": 2, "tStart": 0.0, "DynamZSpAna": True,
    "VIn": 0, "JIn": 0, "MJIn": 0, "EParallel": 0.0, "PhiParallel": 0.0,
    "TolSp": 1.0e-7,
    "writeSMatrix": False, "writeProbs": True, "writeRDProbs": True,
}

SMOKE_DAILIGHT = dict(_COMMON, **{
    "Title": "small Dai-Light H2/Cu(111) grid with angular coupling",
    "SystemID": "H2Cu111DaiLight",
    "ZStart": -1.0, "dZ": 0.16, "nZ": 64, "nZSpecular": 128,
    "RStart": 0.8, "dR": 0.3, "nR": 16,
    "nX": 4, "nY": 4, "JMax": 4, "MJMax": 4,
    "tEnd": 80.0, "dtStep": 2.0, "nStep": 40,
    "dtAna": 10.0, "dtAve": 20.0, "Zana": 7.5, "NrOfEs": 21,
    "EMinWP": 0.2, "EMaxWP": 0.6, "Z0": 11.0,
    "ZOptStart": 8.0, "ZOptEnd": 9.08, "ZOptEnergy": 0.3,
    "ZSpOptStart": 15.0, "ZSpOptEnd": 19.32, "ZSpOptEnergy": 0.3,
    "ROptStart": 4.0, "ROptEnd": 5.3, "ROptEnergy": 0.3,
    "UseSwitching": True, "ZminSwitch": 5.0, "ZmaxSwitch": 7.0,
})

# Shapes without a compile-time kernel instantiation (NZ 72, 6 x 6, nphi 14): the run-time-shaped kernels.
SMOKE_GENERIC = dict(SMOKE_DAILIGHT, **{
    "Title": "small Dai-Light grid on shapes served by the run-time-shaped kernels",
    "dZ": 0.14, "nZ": 72, "nX": 6, "nY": 6, "JMax": 6, "MJMax": 6, "ZOptEnd": 8.94,
})

# The angular bases of the reduced batch members (JMax 12: 7 theta x 28 phi; JMax 18: 10 theta x 40 phi) on the
# small grid.
JMAX12_ANGULAR = dict(SMOKE_DAILIGHT, **{
    "Title": "small Dai-Light grid with JMax 12",
    "JMax": 12, "MJMax": 12, "UseSymmetry": True,
})
JMAX18_ANGULAR = dict(JMAX12_ANGULAR, **{"Title": "small Dai-Light grid with JMax 18", "JMax": 18, "MJMax": 18})

# The angular basis of the DZ workload (JMax 24, MJMax 23: 323 rotational states, 13 theta x 48 phi, C6v symmetry)
# on a small spatial grid: the instantiated MFMA interaction shapes at a size the CPU oracle runs in seconds.
DZ_ANGULAR = dict(SMOKE_DAILIGHT, **{
    "Title": "small Dai-Light grid with the DZ workload's angular basis",
    "JMax": 24, "MJMax": 23, "UseSymmetry": True,
})

# The DZ workload's angular basis and 16 x 16 C6v (x, y) grid on a coarse, short (Z, r) grid: the set-up (Legendre
# matrix at JMax 24, symmetry maps at 16 x 16) at a size whose unreduced potential grid the CPU builds in seconds.
DZ_SETUP = dict(SMOKE_DAILIGHT, **{
    "Title": "coarse Dai-Light grid with the DZ workload's angular basis and (x, y) grid",
    "dZ": 0.32, "nZ": 32, "nZSpecular": 64, "ZOptEnd": 8.92, "ZSpOptEnd": 19.16,
    "dR": 0.6, "nR": 8, "ROptEnd": 5.0,
    "nX": 16, "nY": 16, "JMax": 24, "MJMax": 23, "UseSymmetry": True,
})

# The same with the HardWall workload's angular basis (JMax 10) and an 8 x 8 (x, y) grid.
DZ_SETUP_8X8 = dict(DZ_SETUP, **{
    "Title": "coarse Dai-Light grid with JMax 10 on 8 x 8",
    "nX": 8, "nY": 8, "JMax": 10, "MJMax": 10,
})

HARDWALL_6D = dict(_COMMON, **{
    "Title": "HardWallPES 6D workload",
    "SystemID": "HardWallPES",
    "ZStart": 0.5, "dZ": 19.5 / 127, "nZ": 128, "nZSpecular": 192,
    "RStart": 0.6, "dR": 3.9 / 31, "nR": 32,
    "nX": 8, "nY": 8, "JMax": 10, "MJMax": 10,
    "tEnd": 4000.0, "dtStep": 2.0, "nStep": 2000,
    "dtAna": 10.0, "dtAve": 100.0, "Zana": 10.0, "NrOfEs": 200,
    "EMinWP": 0.05, "EMaxWP": 1.0, "Z0": 14.0,
    "ZOptStart": 16.1, "ZOptEnd": 20.0, "ZOptEnergy": 0.3,
    "ZSpOptStart": 24.0, "ZSpOptEnd": 29.8, "ZSpOptEnergy": 0.3,
    "ROptStart": 3.72, "ROptEnd": 4.5, "ROptEnergy": 0.3,
})

DZ_6D = dict(_COMMON, **{
    "Title": "Dai-Light H2/Cu(111) 6D workload",
    "SystemID": "H2Cu111DaiLight",
    "ZStart": -1.0, "dZ": 13.0 / 95, "nZ": 96, "nZSpecular": 512,
    "RStart": 0.8, "dR": 5.2 / 47, "nR": 48,
    "nX": 16, "nY": 16, "JMax": 24, "MJMax": 23,
    "tEnd": 40000.0, "dtStep": 2.0, "nStep": 20000,
    "dtAna": 10.0, "dtAve": 100.0, "Zana": 8.5, "NrOfEs": 100,
    "EMinWP": 0.2, "EMaxWP": 1.0, "Z0": 14.0,
    "ZOptStart": 9.5, "ZOptEnd": 12.0, "ZOptEnergy": 0.5,
    "ZSpOptStart": 40.0, "ZSpOptEnd": 68.9, "ZSpOptEnergy": 0.5,
    "ROptStart": 4.5, "ROptEnd": 6.0, "ROptEnergy": 0.5,
    "UseSwitching": True, "ZminSwitch": 6.0, "ZmaxSwitch": 8.0,
    "UseSymmetry": True,
})


def _dz_member(vin: int, jin: int, mjin: int, jmax: int | None = None, nr: int = 48) -> dict:
    """A DZ_6D run in complex64 with initial state (v, j, mJ): JMax 24 (MJMax 23) for even j, 23 for odd j (an odd
    j needs an odd JMax for the homonuclear molecule); a given jmax and nr make the reduced members."""
    if jmax is None:
        jmax = 24 if jin % 2 == 0 else 23
    p = dict(DZ_6D)
    p.update({"Precision": "mixed", "VIn": vin, "JIn": jin, "MJIn": mjin, "JMax": jmax, "MJMax": min(jmax, 23),
              "nR": nr, "dR": 5.2 / (nr - 1)})
    return p


# The 12 runs of the batched workload: eight initial states (v, j, mJ) from v in {0, 1} x j in {0 .. 3} x mJ in
# {0 .. j} on the full DZ grid, and four at reduced JMax 12 / 18 with NR 32 (uneven sizes).
BATCH_12 = [
    ("v0j0m0", _dz_member(0, 0, 0)), ("v0j1m0", _dz_member(0, 1, 0)), ("v0j1m1", _dz_member(0, 1, 1)),
    ("v0j2m0", _dz_member(0, 2, 0)), ("v0j2m2", _dz_member(0, 2, 2)), ("v0j3m0", _dz_member(0, 3, 0)),
    ("v1j0m0", _dz_member(1, 0, 0)), ("v1j1m1", _dz_member(1, 1, 1)),
    ("v0j0m0_jmax12_nr32", _dz_member(0, 0, 0, 12, 32)), ("v0j0m0_jmax18_nr32", _dz_member(0, 0, 0, 18, 32)),
    ("v1j0m0_jmax12_nr32", _dz_member(1, 0, 0, 12, 32)), ("v0j2m1_jmax18_nr32", _dz_member(0, 2, 1, 18, 32)),
]


def write_case(name: str, params: dict, directory: str | None = None) -> str:
    """Writes params as <directory>/<name>.json (a temporary directory by default) and returns the path."""
    directory = directory or os.path.join(tempfile.gettempdir(), "spodvr_cases")
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, f"{name}.json")
    with open(path, "w") as f:
        json.dump(params, f, indent=1)
    return path


def golden_input(case: str) -> str:
    return os.path.join(GOLDEN_DIR, case, "input.json")


def golden_overrides(case: str) -> dict:
    """The data files of a golden case that names them (the CRP data set, the SCM parameter and slab files), as the
    fixture copies under tests/golden/."""
    out = {}
    doc = json.load(open(golden_input(case)))
    if doc.get("SystemID") in ("CRPlibrary", "CRPDistrtSurface"):
        out["CRPlibDataSet"] = os.path.join(GOLDEN_DIR, "crp", "cu111_srp48", "dataset")
    if doc.get("SystemID") == "CRPDistrtSurface":
        out["SCMParFilepath"] = os.path.join(GOLDEN_DIR, "scm", "example", "params.dat")
        out["PosFilepath"] = os.path.join(GOLDEN_DIR, "scm", "example", "examplePosfile.dat")
    return out
