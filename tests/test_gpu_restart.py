"""Restarts on the GPU path: a run resumed from tracer_finalize / tracer_init equals the uninterrupted run bit
for bit (the restart carries every array the step reads), and tracer_run_workflow (init, steps of tracer 1,
finalize into the same directory) leaves a restart from which tracer 1 continues as if never stopped."""
import numpy as np
import pytest

from test_restart import _setup

pytestmark = pytest.mark.gpu


def test_resume_from_restart_is_bitwise(api, tmp_path):
    print("[gpu] restart resume", flush=True)
    _setup(api)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    api.advect_steps(3, dt)
    straight = [api.get_tracer_values(i + 1) for i in range(3)]
    _setup(api)
    api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    assert api.finalize(str(tmp_path / "r")) == 0
    assert api.initialize(str(tmp_path / "r")) == 0
    api.advect_steps(3, dt)
    for i in range(3):
        assert np.array_equal(api.get_tracer_values(i + 1), straight[i]), f"tracer {i + 1}"


def test_run_workflow_continues(api, tmp_path):
    print("[gpu] tracer_run_workflow", flush=True)
    _setup(api)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    d = tmp_path / "wf"
    assert api.finalize(str(d)) == 0
    assert api.initialize(str(d)) == 0
    for _ in range(4):
        assert api.advect_step(dt) == 0
    ref = api.get_tracer_values(1)
    assert api.finalize(str(tmp_path / "unused")) == 0
    assert api.run_workflow(str(d), 2, dt) == 0
    assert api.run_workflow(str(d), 2, dt) == 0
    assert api.initialize(str(d)) == 0
    assert np.array_equal(api.get_tracer_values(1), ref)
