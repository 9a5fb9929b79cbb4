"""Readers of the golden fixtures (tests/golden/<case>/, extracted by tests/golden/make_fixtures.py from the
reference's Fortran outputs) and the gate comparisons of the reference's regression tests:
  averages   every printed column within half a unit of its last printed digit
             (tests/propagation_reference.hpp:285-306 of the reference);
  BK         every coefficient of every analysis step within 1e-10 * max(max |c|, 1e-4)
             (tests/regression_gate.hpp:19, tests/analysis_reference.hpp:556-580);
  fort.25    S within its printed digits plus 1e-10 * max(max |S|, 1e-3), P within half its printed unit plus the
             effect of that difference, the phase where that bound is at most 1e-3
             (tests/fort25_reference.hpp:64-199)."""

from __future__ import annotations

import json
import math
import os
import re

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case_dir(case: str) -> str:
    return os.path.join(GOLDEN, case)


def fortran_float(text: str) -> float:
    return float(text.replace("D", "E").replace("d", "e"))


def half_unit(text: str, floor: float = 0.0) -> float:
    """Half a unit in the last printed digit of a Fortran number (with a 1% margin), at least floor; an
    E/D-format zero is exact."""
    t = text.replace("D", "E").replace("d", "e")
    m = re.fullmatch(r"([-+]?\d*)\.(\d*)(?:E([-+]?\d+))?", t)
    if not m:
        return float("nan")
    mantissa = m.group(1) + m.group(2)
    if m.group(3) is not None and not re.search(r"[1-9]", mantissa):
        return floor
    digits = len(m.group(2))
    exponent = int(m.group(3) or 0)
    return max(1.01 * 0.5 * 10.0 ** (exponent - digits), floor)


def read_averages(case: str) -> list[list[str]]:
    rows = []
    for line in open(os.path.join(case_dir(case), "averages.out")):
        parts = line.split()
        if len(parts) == 5 and all(re.fullmatch(r"-?\d\.\d+E[-+]\d+", p) for p in parts):
            rows.append(parts)
    return rows


def compare_averages(ours: np.ndarray, case: str) -> float:
    """Largest |ours - Fortran| in units of the printed half unit (must be <= 1); rows matched by count."""
    ref = read_averages(case)
    assert len(ours) == len(ref), f"{len(ours)} averages rows, Fortran {len(ref)}"
    worst = 0.0
    for mine, theirs in zip(ours, ref):
        for a, text in zip(mine, theirs):
            half = half_unit(text)
            diff = abs(a - fortran_float(text))
            if half == 0.0:  # a printed zero is exact
                worst = max(worst, 0.0 if diff == 0.0 else math.inf)
            else:
                worst = max(worst, diff / half)
    return worst


def load_bk(case: str):
    return np.load(os.path.join(case_dir(case), "bk.npz"))


def compare_bk(records: np.ndarray, case: str) -> tuple[float, float]:
    """(largest |ours - Fortran| over the kept coefficients of every record, the gate scale max(max|c|, 1e-4))."""
    f = load_bk(case)
    assert records.shape == (int(f["n_records"]), int(f["n_coefficients"])), records.shape
    diff = np.abs(records[:, f["indices"]] - f["records"]).max()
    return float(diff), max(float(f["scale"]), 1e-4)


def load_stdout(case: str) -> dict:
    return json.load(open(os.path.join(case_dir(case), "stdout.json")))


def compare_fort25(run, results: dict, case: str, code_fraction: float = 1e-10) -> dict:
    """Checks the kept fort.25 rows against run.results(); returns counts and the worst code difference."""
    f = np.load(os.path.join(case_dir(case), "fort25.npz"))
    state = {(int(j), int(mj), int(n), int(m)): k for k, (j, mj, n, m) in enumerate(
        zip(run.array("state.j"), run.array("state.mj"), run.array("state.n"), run.array("state.m")))}
    order = run.array("state.order")
    n_states = run.info.n_states
    s_scale = max(abs(complex(fortran_float(re_), fortran_float(im))) for re_, im, _, _ in f["values"])
    scale = max(s_scale, 1e-3)
    code = code_fraction * scale
    out = {"rows": 0, "failed": 0, "phase_compared": 0, "worst": 0.0, "failures": []}
    energies = results["total_energy"]
    for b in f["blocks"]:
        text = str(f["total_energy"][b])
        if abs(energies[b] - fortran_float(text)) > half_unit(text):
            out["failed"] += 1
            out["failures"].append(f"block {b}: total energy {energies[b]} vs {text}")
    for (block, v, j, mj, od, n, m), (re_, im, p, ph) in zip(f["rows"], f["values"]):
        out["rows"] += 1
        k = state.get((int(j), int(mj), int(n), int(m)))
        if k is None or int(order[k]) != int(od):
            out["failed"] += 1
            out["failures"].append(f"no state for row {(block, v, j, mj, od, n, m)}")
            continue
        idx = k + n_states * int(v)
        s_c = results["s"][block, idx]
        s_f = complex(fortran_float(re_), fortran_float(im))
        digits = math.hypot(half_unit(re_), half_unit(im))
        err = abs(s_c - s_f)
        if err > digits:
            out["worst"] = max(out["worst"], (err - digits) / scale)
        s_abs = abs(s_f)
        p_tol = half_unit(p, 1e-20) + 2.0 * (s_abs + digits) * code + code * code
        ok = err <= digits + code and abs(results["p"][block, idx] - fortran_float(p)) <= p_tol
        phase_c = math.atan2(s_c.imag, s_c.real) if s_c.real != 0.0 else 0.0
        if fortran_float(re_) != 0.0 and s_abs > 0.0:
            ptol = half_unit(ph) + code / s_abs
            if ptol <= 1e-3:
                out["phase_compared"] += 1
                d = math.remainder(phase_c - fortran_float(ph), 2 * math.pi)
                ok = ok and abs(d) <= ptol
        if not ok:
            out["failed"] += 1
            if len(out["failures"]) < 10:
                out["failures"].append(f"row {(block, v, j, mj, n, m)}: S {s_c} vs {s_f}")
    return out


def full_run(run) -> dict:
    """Runs the whole time loop of a case and collects what the gates compare."""
    run.steps(10 ** 9)
    return {"averages": run.averages(), "bk": run.bk_records(), "results": run.results(),
            "surface": run.surface(), "limits": run.array("analysis.limits")}


def check_gates(run, out: dict, case: str) -> dict:
    """Gates 3 and 4 of the reference for one case; returns the measured margins (asserts inside)."""
    report = {}
    report["averages_half_units"] = compare_averages(out["averages"], case)
    assert report["averages_half_units"] <= 1.0, report
    diff, scale = compare_bk(out["bk"], case)
    report["bk_relative"] = diff / scale
    assert diff <= 1e-10 * scale, report
    stdout = load_stdout(case)
    names = ["jana", "nviba", "nana", "nstata", "inchna"]
    for name, mine in zip(names, out["limits"]):
        assert stdout["limits"][name] == [int(mine)], (name, stdout["limits"][name], mine)
    fixed, _, z_line = out["surface"]
    assert fixed == bool(stdout["fixed_line"]), (fixed, stdout["fixed_line"])
    if fixed:
        text = stdout["fixed_line"][0]
        assert abs(z_line - fortran_float(text)) <= half_unit(text), (z_line, text)
    f25 = compare_fort25(run, out["results"], case)
    report["fort25_rows"] = f25["rows"]
    report["fort25_phase_compared"] = f25["phase_compared"]
    report["fort25_worst_code_difference"] = f25["worst"]
    assert f25["failed"] == 0, f25["failures"]
    return report
