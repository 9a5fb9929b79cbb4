"""The calls the time loop makes across the backend seam, counted through the decorator every run is made behind
(spo_run_call_counts), after the reference's tests/test_time_loop_calls.cpp:29-113 with its CountingBackend
(tests/counting_backend.hpp:35-113). The backend here is the CPU oracle; the HIP backend sits behind the same
decorator and driver.

This driver's time loop moves the Balint-Kurti projection below the seam: per analysis step it makes one
propagate_analysis_step, one analyse (the coefficients, not the analysis plane) and one specular read (the dynamic
analysis surface), and no z_plane; the block densities once per averages row (tStart and every dtAve); the whole wave
packet on the host only for a checkpoint, after the loop; a restart reads the wave packet once and hands the earlier
run's records to the sums one call each. The input is the reference's small hard-wall case of these tests
(hardwall_short on a 2 x 2 (x, y) grid, JMax = MJMax = 2, tEnd 1000 a.u.: 100 analysis steps, 11 averages rows)."""

import json
import os

import pytest

from tests import cases

HERE = os.path.dirname(os.path.abspath(__file__))
ANALYSIS_STEPS = 100
AVERAGES_ROWS = 11  # tStart and every dtAve


def _small_hardwall(**extra):
    with open(os.path.join(HERE, "golden", "hardwall_short", "input.json")) as f:
        params = json.load(f)
    params.update({"nX": 2, "nY": 2, "JMax": 2, "MJMax": 2, "tEnd": 1000.0, "nStep": 500})
    params.update(extra)
    return params


def _oracle_run(name, params, overrides=None):
    from oracle import oracle
    return oracle.run(cases.write_case(name, params), overrides)


def _expect_base(c, steps=ANALYSIS_STEPS, rows=AVERAGES_ROWS):
    assert c["propagate_analysis_step"] == steps
    assert c["analyse"] == steps
    assert c["specular"] == steps + rows  # the analysis surface every step, the averages every row
    assert c["block_densities"] == rows
    assert c["z_plane"] == 0
    assert c["kinetic_and_potential"] == 0
    assert c["read_wave_packet"] == 0
    assert c["site_column"] == 0
    assert c["add_earlier_analysis_step"] == 0
    assert c["energy_resolved_add"] == 0


def test_time_loop_calls_without_extra_keys():
    run = _oracle_run("seam_calls", _small_hardwall())
    try:
        assert run.info.analysis_steps_total == ANALYSIS_STEPS
        run.steps(ANALYSIS_STEPS)
        c = run.call_counts()
        _expect_base(c)
        assert c["wave_packet_on_host"] == 0
        assert c["s_matrix_sums"] == 0
        run.results()  # the S matrix from the device sums: one read of the sums
        assert run.call_counts()["s_matrix_sums"] == 1
    finally:
        run.close()


def test_checkpoint_reads_the_wave_packet_once_after_the_loop(tmp_path):
    run = _oracle_run("seam_calls_ckpt", _small_hardwall())
    try:
        run.steps(ANALYSIS_STEPS)
        run.write_checkpoint(str(tmp_path))
        c = run.call_counts()
        _expect_base(c)
        assert c["wave_packet_on_host"] == 1
    finally:
        run.close()


def test_restart_reads_the_wave_packet_once_and_replays_the_records(tmp_path):
    first = _oracle_run("seam_calls_first", _small_hardwall(tEnd=500.0, nStep=250))
    try:
        first.steps(50)
        first.write_checkpoint(str(tmp_path))
    finally:
        first.close()
    second = _oracle_run("seam_calls_second", _small_hardwall(NewCalc=False, RestartDir=str(tmp_path), tStart=500.0,
                                                             tEnd=1000.0, nStep=250))
    try:
        c = second.call_counts()
        assert c["read_wave_packet"] == 1
        assert c["add_earlier_analysis_step"] == 50
        second.steps(50)
        c = second.call_counts()
        assert c["propagate_analysis_step"] == 50 and c["analyse"] == 50 and c["z_plane"] == 0
        assert c["wave_packet_on_host"] == 0 and c["site_column"] == 0
    finally:
        second.close()


def test_energy_resolved_sums_add_one_call_per_step_and_read_no_column():
    params = _small_hardwall(PlotSites=[{"name": "a", "x": 0.0, "y": 0.0}, {"name": "b", "x": 0.5, "y": 0.0}],
                             EnergyResolvedSites=["a", "b"], EnergyResolvedEnergies=[0.3, 0.4, 0.5])
    run = _oracle_run("seam_calls_er", params)
    try:
        run.steps(ANALYSIS_STEPS)
        c = run.call_counts()
        assert c["energy_resolved_add"] == ANALYSIS_STEPS  # both sites in one call: the columns stay below the seam
        assert c["site_column"] == 0
        assert c["propagate_analysis_step"] == ANALYSIS_STEPS and c["analyse"] == ANALYSIS_STEPS
        assert c["block_densities"] == AVERAGES_ROWS and c["wave_packet_on_host"] == 0
    finally:
        run.close()


def test_batched_steps_are_counted_per_run():
    import spodvr_amd
    runs = [_oracle_run(f"seam_calls_batch{k}", _small_hardwall()) for k in range(2)]
    try:
        spodvr_amd.run_batch(runs, 20)
        for r in runs:
            c = r.call_counts()
            assert c["propagate_analysis_step"] == 20 and c["analyse"] == 20
    finally:
        for r in runs:
            r.close()


def test_counts_are_planted_mistake_sensitive():
    """A stage-level call shows up in its own counter and nowhere else."""
    run = _oracle_run("seam_calls_stage", _small_hardwall())
    try:
        before = run.call_counts()
        run.kinetic_and_potential(True, False)
        run.block_densities()
        after = run.call_counts()
        changed = {k for k in after if after[k] != before[k]}
        assert changed == {"kinetic_and_potential", "block_densities"}
        assert after["kinetic_and_potential"] == before["kinetic_and_potential"] + 1
    finally:
        run.close()


@pytest.mark.gpu
def test_hip_backend_makes_the_same_calls():
    import spodvr_amd
    run = spodvr_amd.Run(cases.write_case("seam_calls_hip", _small_hardwall()), device="hip")
    try:
        run.steps(ANALYSIS_STEPS)
        _expect_base(run.call_counts())
    finally:
        run.close()
