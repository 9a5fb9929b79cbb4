"""Graph mode (SPO_HIP_GRAPH=1 -> spo_hip_graph_enable: one analysis step captured into a hipGraph, then replayed)
against the same run stepped eagerly: after every analysis step (each a replay, with the run's analysis and
averages launched eagerly between the replays) the BK record, the S-matrix sums and the averages are bitwise the
eager run's, in both precisions and in a batch; and the replays' time per step. Every test prints its case first."""

import os
import sys
import time

import numpy as np
import pytest

import spodvr_amd
from tests import cases

pytestmark = pytest.mark.gpu


def _note(*what):
    print("graph:", *what, file=sys.stderr, flush=True)


def _run(case, precision, graph):
    old = os.environ.get("SPO_HIP_GRAPH")
    os.environ["SPO_HIP_GRAPH"] = "1" if graph else "0"
    try:
        return spodvr_amd.Run(cases.golden_input(case), {"Precision": precision}, device="hip")
    finally:
        if old is None:
            os.environ.pop("SPO_HIP_GRAPH", None)
        else:
            os.environ["SPO_HIP_GRAPH"] = old


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_replays_equal_eager_steps(precision):
    _note("eager vs graph, cu111_dailight_short", precision)
    eager = _run("cu111_dailight_short", precision, False)
    graph = _run("cu111_dailight_short", precision, True)
    try:
        for step in range(6):
            eager.steps(1)
            graph.steps(1)
            assert np.array_equal(graph.bk_records(), eager.bk_records()), step
            assert np.array_equal(graph.s_sums(), eager.s_sums()), step
        assert np.array_equal(np.array(graph.averages()), np.array(eager.averages()))
    finally:
        eager.close()
        graph.close()


def test_graph_runs_in_a_batch():
    _note("batch of graph and eager runs, hardwall_short")
    single = _run("hardwall_short", "mixed", False)
    try:
        single.steps(5)
        ref = single.bk_records()
    finally:
        single.close()
    runs = [_run("hardwall_short", "mixed", True), _run("hardwall_short", "mixed", False)]
    try:
        spodvr_amd.run_batch(runs, 5)
        for r in runs:
            assert np.array_equal(r.bk_records(), ref)
    finally:
        for r in runs:
            r.close()


def test_replay_time():
    for graph in (False, True):
        run = _run("hardwall_short", "mixed", graph)
        try:
            run.steps(2)
            t0 = time.perf_counter()
            run.steps(40)
            dt = time.perf_counter() - t0
            _note(f"hardwall_short mixed, graph {graph}: {dt / 40 * 1e3:.3f} ms per analysis step")
        finally:
            run.close()
