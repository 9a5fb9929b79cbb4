"""The HardWall workload (NZ 128 x NR 32 x 8 x 8, JMax 10, 2000 time steps of 2 a.u.) run to its end from its real
initial wave packet: the FP64 GPU run against the CPU oracle on the same input (the probabilities within 5e-7 of
the largest one, the reference's seventh-digit criterion for its CUDA path), and the complex64 run against the FP64
GPU run (|dP_reaction| <= 1e-4, the benchmark metric's condition). The oracle's run takes about a minute on the
host's cores."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def results():
    import spodvr_amd
    from oracle import oracle
    path = cases.write_case("hardwall_6d", cases.HARDWALL_6D)
    out = {}
    for name, precision in (("fp64", "fp64"), ("mixed", "mixed")):
        print(f"hardwall_6d {name}: 2000 time steps on the GPU", flush=True)
        run = spodvr_amd.Run(path, {"Precision": precision}, device="hip")
        run.steps(10 ** 9)
        out[name] = run.results()
        run.close()
    print("hardwall_6d: 2000 time steps on the CPU oracle", flush=True)
    cpu = oracle.run(path)
    cpu.steps(10 ** 9)
    out["oracle"] = cpu.results()
    cpu.close()
    return out


def test_fp64_probabilities_match_the_oracle(results):
    a, b = results["oracle"], results["fp64"]
    scale = float(np.abs(a["p"]).max())
    dp = float(np.abs(a["p"] - b["p"]).max())
    dr = float(np.abs(a["p_reaction"] - b["p_reaction"]).max())
    print(f"  FP64 GPU vs oracle: max |dP| {dp:.3e}, max |dP_reaction| {dr:.3e} (largest P {scale:.4f})", flush=True)
    assert dp <= 5e-7 * scale and dr <= 5e-7 * max(scale, float(np.abs(a["p_reaction"]).max()))


def test_mixed_reaction_probability_within_1e4_of_fp64(results):
    a, b = results["fp64"], results["mixed"]
    dp = float(np.abs(a["p"] - b["p"]).max())
    dr = float(np.abs(a["p_reaction"] - b["p_reaction"]).max())
    print(f"  complex64 vs FP64: max |dP| {dp:.3e}, max |dP_reaction| {dr:.3e}", flush=True)
    assert np.all(np.isfinite(b["p_reaction"]))
    assert dr <= 1e-4 and dp > 0.0
