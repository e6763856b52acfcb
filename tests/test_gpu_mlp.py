"""The cost model's GPU kernels against the float64 oracle, and the learned
planner's plans against the state oracle."""
import os

import numpy as np
import pytest

from ket_amd import costmodel as CM
from ket_amd import native as N
from ket_amd import workloads as W
from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_forward_and_gradients_match_oracle():
    rng = np.random.default_rng(3)
    mlp = N.MLP(seed=5)
    p = mlp.params.astype(np.float64)
    x = rng.normal(size=(300, N.MLP_INPUTS)).astype(np.float32)
    t = rng.normal(size=300).astype(np.float32)
    y = mlp.forward(x)
    want = M.forward(p, x)
    assert np.abs(y - want).max() <= 1e-4 * max(1.0, np.abs(want).max())
    g, loss = mlp.gradients(x, t)
    wl, wg = M.loss_and_grad(p, x, t)
    assert abs(loss - wl) <= 1e-4 * max(1.0, wl)
    assert np.linalg.norm(g - wg) <= 1e-4 * np.linalg.norm(wg)


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, size=(4096, N.MLP_INPUTS)).astype(np.float32)
    t = (np.sin(3 * x[:, 0]) + x[:, 1] * x[:, 2] - 0.5 * x[:, 3]).astype(np.float32)
    mlp = N.MLP(seed=1)
    first = mlp.train_step(x, t, 3e-3)
    for _ in range(400):
        last = mlp.train_step(x, t, 3e-3)
    assert last < first / 10


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_learned_plans_keep_the_state(seed):
    mlp = CM.load() or N.MLP(seed)
    mlp.use_for_planning(True)
    try:
        for c in (W.random_mixed(12, 12, seed), W.qft_u(11, seed), W.diag_mix(10, 5, seed)):
            want = O.run(c)
            with N.StateVector(c.n_qubits, N.C128) as sv:
                sv.apply_circuit(c, N.FUSE_LEARNED, 5)
                got = sv.to_numpy()
            assert np.abs(got - want).max() <= 1e-12
    finally:
        mlp.use_for_planning(False)


@pytest.mark.skipif(not os.path.exists(CM.MODEL_PATH), reason="no trained cost model committed")
def test_trained_planner_against_fixed_forms():
    """The trained planner on held-out circuits of the training distribution
    (26-27 qubits here, complex64, resident states), against every fixed
    form: whole-circuit wall time, each form warmed and timed as the median
    of 3 cached repeats. It must be no slower than the best non-learned form
    beyond the measured timing noise, and faster than the method without the
    network (ket's Fuser as exact runs, its fallback) by more than it."""
    mlp = CM.load()
    states = CM.States(N.C64)
    try:
        circs = CM.config5_circuits(6, seed=999, lo=26, hi=27, depth=40)
        res = CM.evaluate(mlp, circs, states, reps=3)
    finally:
        states.close()
    print("learned planner vs fixed forms:", res)
    tot, noise = res["total_ms"], res["noise"]
    assert tot["learned"] <= tot[res["best_fixed_form"]] * (1 + noise)
    assert tot["learned"] * (1 + noise) < tot[CM.FALLBACK]


@pytest.mark.skipif(not os.path.exists(CM.MODEL_PATH), reason="no trained cost model committed")
def test_trained_planner_takes_blocks_on_phase_ladders():
    """Where the forms trade places: on a 30-qubit QFT (after a random U
    layer, complex64) k = 5 blocks beat the run packer, on 30-qubit
    brickwork the run packer beats blocks. The learned planner must be as
    fast as the fastest fixed form on both (within 3%), so it beats the run
    packer on the QFT and the block packer on the brickwork."""
    mlp = CM.load()
    states = CM.States(N.C64)
    try:
        res = CM.time_cases(mlp, [("qft", W.qft_u(30, seed=3)),
                                  ("brick", W.brickwork_mixed(30, 24, seed=4))], states)
    finally:
        states.close()
    print("learned planner on QFT and brickwork:", res)
    for name, row in res.items():
        best = min(v["ms"] for f, v in row.items() if f != "learned")
        assert row["learned"]["ms"] <= best * 1.03, name
    assert res["qft"]["learned"]["ms"] < 0.95 * res["qft"]["packed_runs"]["ms"]
    assert res["brick"]["learned"]["ms"] < 0.95 * res["brick"]["greedy_blocks"]["ms"]
