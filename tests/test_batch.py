"""Runs in lockstep (spo_run_steps_batch, Propagation::run_batch): each run's BK records, S-matrix sums and
averages equal those of the same run stepped alone, bit for bit, also when the runs have different lengths.
Here with the CPU oracle's backend (propagate_batch steps non-HIP backends one after another); the GPU test
(tests/test_gpu_batch.py) runs the same on the HIP backend, whose steps are enqueued together."""

import numpy as np

import spodvr_amd
from oracle import oracle
from tests import cases


def _results(run):
    return run.bk_records(), run.s_sums(), np.array(run.averages())


def test_batch_equals_single_runs():
    path = cases.golden_input("hardwall_short")
    single = []
    for steps in (3, 2):
        run = oracle.run(path)
        run.steps(steps)
        single.append(_results(run))
        run.close()
    runs = [oracle.run(path), oracle.run(path)]
    spodvr_amd.run_batch(runs[1:], 1)  # the second run one step ahead,
    spodvr_amd.run_batch(runs, 2)      # then both in lockstep: 2 and 3 steps in all
    batched = [_results(runs[0]), _results(runs[1])]
    for r in runs:
        r.close()
    for (bk, s, av), (bk_ref, s_ref, av_ref) in zip(batched, [single[1], single[0]]):
        assert bk.shape == bk_ref.shape
        assert np.array_equal(bk, bk_ref) and np.array_equal(s, s_ref) and np.array_equal(av, av_ref)
