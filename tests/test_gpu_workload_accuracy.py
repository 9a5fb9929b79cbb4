"""The benchmark metric's accuracy condition at its own workload: the Dai-Light H2/Cu(111) 6D run (NZ 96 x NR 48,
16 x 16, JMax 24 / MJMax 23, C6v, specular grid 512) propagated from its real initial wave packet for its configured
20000 time steps (4000 analysis steps of 5), once with an FP64 wave packet and once with the complex64 storage mode,
on the GPU. The reaction probability P_reaction(E) = 1 - sum of the state-to-state P (src/analysis/s_matrix.cpp:
118-176) of the mixed run must be within 1e-4 of the FP64 run's at every analysed energy; the largest difference of a
state-to-state probability (v, j, mJ, n, m) is reported beside it.

The runs are advanced in segments of about a minute each (one test per segment), so that progress shows while they
run; the last test compares them. The FP64 path itself is held to the CPU oracle step by step at this shape
(test_gpu_parity_configs.py) and end to end by the Fortran gates on the smaller gated cases."""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

SEGMENTS_FP64 = 4
SEGMENTS_MIXED = 2


class Pair:
    def __init__(self):
        import spodvr_amd
        self.path = cases.write_case("dz_6d", cases.DZ_6D)
        print("dz_6d: FP64 and complex64 runs, real initial wave packet", flush=True)
        self.fp64 = spodvr_amd.Run(self.path, device="hip")
        self.mixed = spodvr_amd.Run(self.path, {"Precision": "mixed"}, device="hip")
        self.total = self.fp64.info.analysis_steps_total
        assert self.total == self.mixed.info.analysis_steps_total
        self.done = {"fp64": 0, "mixed": 0}

    def advance(self, which, segments, k):
        run = self.fp64 if which == "fp64" else self.mixed
        end = self.total * (k + 1) // segments
        n = end - self.done[which]
        print(f"dz_6d {which}: analysis steps {self.done[which]} .. {end} of {self.total}", flush=True)
        run.steps(n)
        self.done[which] = end

    def close(self):
        self.fp64.close()
        self.mixed.close()


@pytest.fixture(scope="module")
def pair():
    p = Pair()
    yield p
    p.close()


@pytest.mark.parametrize("k", range(SEGMENTS_FP64))
def test_fp64_segment(pair, k):
    pair.advance("fp64", SEGMENTS_FP64, k)


@pytest.mark.parametrize("k", range(SEGMENTS_MIXED))
def test_mixed_segment(pair, k):
    pair.advance("mixed", SEGMENTS_MIXED, k)


def test_mixed_reaction_probability_within_1e4_of_fp64(pair):
    assert pair.done["fp64"] == pair.total and pair.done["mixed"] == pair.total
    a = pair.fp64.results()
    b = pair.mixed.results()
    energies = a["total_energy"]
    dpr = np.abs(a["p_reaction"] - b["p_reaction"])
    dp = np.abs(a["p"] - b["p"])
    worst_e = int(np.argmax(dpr))
    print(f"dz_6d after {pair.total} analysis steps: P_reaction(FP64) in [{a['p_reaction'].min():.6f}, "
          f"{a['p_reaction'].max():.6f}] over {len(energies)} energies ({energies[0]:.4f} .. {energies[-1]:.4f} eV)",
          flush=True)
    print(f"  max |dP_reaction| {dpr.max():.3e} at E = {energies[worst_e]:.4f} eV; "
          f"max |dP| over (v, j, mJ, n, m) {dp.max():.3e}; sum P(FP64) max {a['p'].sum(axis=1).max():.6f}",
          flush=True)
    fa, fb = pair.fp64.averages(), pair.mixed.averages()
    print(f"  final norm FP64 {fa[-1, 1]:.8f} mixed {fb[-1, 1]:.8f}", flush=True)
    assert np.all(np.isfinite(a["p_reaction"])) and np.all(np.isfinite(b["p_reaction"]))
    assert dpr.max() <= 1e-4
