"""Runs in lockstep on the GPU (spo_run_steps_batch -> spo_hip_propagate_batch: the runs' time steps enqueued
together on their own streams): every run's BK records and S-matrix sums equal those of the run stepped alone, bit
for bit (the same kernels in the same order per run), for runs of different sizes and precisions; and a batch of
small runs takes less time than the same runs one after another. Every test prints its case first."""

import sys
import time

import numpy as np
import pytest

import spodvr_amd
from tests import cases

pytestmark = pytest.mark.gpu


def _note(*what):
    print("batch:", *what, file=sys.stderr, flush=True)


def _run(case, precision):
    return spodvr_amd.Run(cases.golden_input(case), {"Precision": precision}, device="hip")


MEMBERS = [("hardwall_short", "fp64"), ("cu111_dailight_short", "fp64"), ("hardwall_short", "mixed")]


def test_batch_equals_single_runs():
    steps = 4
    single = []
    for case, precision in MEMBERS:
        _note("single", case, precision)
        run = _run(case, precision)
        try:
            run.steps(steps)
            single.append((run.bk_records(), run.s_sums()))
        finally:
            run.close()
    _note("batch of", len(MEMBERS))
    runs = [_run(c, p) for c, p in MEMBERS]
    try:
        spodvr_amd.run_batch(runs, steps)
        for (bk, s), run in zip(single, runs):
            assert np.array_equal(run.bk_records(), bk)
            assert np.array_equal(run.s_sums(), s)
    finally:
        for r in runs:
            r.close()


def test_batch_of_small_runs_is_faster():
    n_runs, steps = 6, 20
    runs = [_run("hardwall_short", "mixed") for _ in range(n_runs)]
    try:
        for r in runs:  # warm-up (launch variants, first launches)
            r.steps(1)
        _note("sequential", n_runs, "runs x", steps, "analysis steps")
        t0 = time.perf_counter()
        for r in runs:
            r.steps(steps)
        sequential = time.perf_counter() - t0
        _note("batched")
        t0 = time.perf_counter()
        spodvr_amd.run_batch(runs, steps)
        batched = time.perf_counter() - t0
        _note(f"sequential {sequential:.3f} s, batched {batched:.3f} s ({sequential / batched:.2f}x)")
        assert batched < sequential
    finally:
        for r in runs:
            r.close()


def _graph_mode(on):
    import os
    if on:
        os.environ["SPO_HIP_GRAPH"] = "1"
    else:
        os.environ.pop("SPO_HIP_GRAPH", None)


def test_batch_graph_equals_single_runs():
    """The batch in graph mode (SPO_HIP_GRAPH=1: one hipGraph captured over every member's streams, one launch per
    analysis step, spo_hip_batch_*) gives each member the BK records and S sums of the member stepped alone, bit for
    bit, over several replays; the members differ in size, precision and steps per analysis."""
    steps = 4
    single = []
    for case, precision in MEMBERS:
        _note("single (eager)", case, precision)
        run = _run(case, precision)
        try:
            run.steps(steps)
            single.append((run.bk_records(), run.s_sums()))
        finally:
            run.close()
    _graph_mode(True)
    try:
        _note("graph batch of", len(MEMBERS))
        runs = [_run(c, p) for c, p in MEMBERS]
    finally:
        _graph_mode(False)
    try:
        _graph_mode(True)
        try:
            spodvr_amd.run_batch(runs, steps)
        finally:
            _graph_mode(False)
        for (bk, s), run in zip(single, runs):
            assert np.array_equal(run.bk_records(), bk)
            assert np.array_equal(run.s_sums(), s)
            assert run.call_counts()["propagate_analysis_step"] == steps
    finally:
        for r in runs:
            r.close()


# Three members of the 12-run DZ batch: a full one and the two reduced-JMax, NR 32 ones (other rotational bases,
# phi steps of 28 and 40 points, another r grid), as the batch workload runs them (complex64, graph mode).
DZ_MEMBERS = [cases.BATCH_12[0], cases.BATCH_12[8], cases.BATCH_12[9]]


def test_dz_batch_members_graph_equals_single_runs():
    """Uneven members of the DZ batch workload in one graph-mode batch: each member's BK records and S sums equal
    those of the member stepped alone (eagerly), bit for bit."""
    steps = 2
    single = []
    for name, params in DZ_MEMBERS:
        _note("single (eager)", name)
        run = spodvr_amd.Run(cases.write_case(f"batch_{name}", params), device="hip")
        try:
            run.steps(steps)
            single.append((run.bk_records(), run.s_sums()))
        finally:
            run.close()
    _graph_mode(True)
    try:
        _note("graph batch of", [name for name, _ in DZ_MEMBERS])
        runs = [spodvr_amd.Run(cases.write_case(f"batch_{name}", params), device="hip") for name, params in DZ_MEMBERS]
    finally:
        _graph_mode(False)
    try:
        _graph_mode(True)
        try:
            spodvr_amd.run_batch(runs, steps)
        finally:
            _graph_mode(False)
        for (bk, s), run in zip(single, runs):
            assert np.array_equal(run.bk_records(), bk)
            assert np.array_equal(run.s_sums(), s)
    finally:
        for r in runs:
            r.close()
