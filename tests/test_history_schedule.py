"""The output-driven step control of the reference's main loop
(src/main.cpp:155-345, src/main_triggers.hpp:26-130) as z4c_amd.awa.HistorySchedule
mirrors it: steps are shortened to land on each history time, rows are
written at the accumulated output times, dt returns to the block dt, and
the run ends exactly at tlim."""
import numpy as np

from z4c_amd.awa import HistorySchedule


def _run(block_dt, dt_hst, tlim):
    s = HistorySchedule(block_dt, dt_hst, tlim)
    t0, _ = s.initial_row()
    its = list(s.iterations())
    return t0, its


def test_rows_land_on_output_times_gauge_wave_cadence():
    """dt = 1/256 (64 cells, CFL 0.25), rows every 0.1 to t = 10: 0.1 is
    25.6 steps, so each interval takes 25 full steps and one shortened step
    that lands on the output time; the output times are the running sums
    0.1 + 0.1 + ... (history.cpp:451), so the last interval ends at
    9.99999999999998 and a clamped step of tlim - time follows."""
    t0, its = _run(1.0 / 256, 0.1, 10.0)
    assert t0 == 0.0
    rows = [it for it in its if it["row"]]
    expect, acc = [], 0.0
    for _ in range(100):
        acc += 0.1
        expect.append(acc)
    expect.append(10.0)
    assert [r["time"] for r in rows] == expect  # exact, not approximate
    assert all(r["constraints"] for r in rows)
    dts = np.array([it["dt"] for it in its])
    assert len(its) == 2601
    assert np.all(dts <= 1.0 / 256)
    full = np.isclose(dts, 1.0 / 256, rtol=0, atol=0)
    assert full.sum() == 2500
    # every step that is not a full step ends on an output time or on tlim
    ends = {r["time"] for r in rows}
    assert all(it["time"] in ends for it, f in zip(its, full) if not f)
    # time accumulates to tlim exactly
    t = 0.0
    for d in dts:
        t += d
    assert t == 10.0


def test_commensurate_cadence_needs_no_shortening():
    """dt_hst a multiple of the block dt: no step is shortened and rows come
    every 4 steps."""
    _, its = _run(1.0 / 16, 0.25, 1.0)
    assert all(it["dt"] == 1.0 / 16 for it in its)
    assert [i for i, it in enumerate(its) if it["row"]] == [3, 7, 11, 15]
    assert [it["time"] for it in its if it["row"]] == [0.25, 0.5, 0.75, 1.0]


def test_first_iteration_evaluates_the_constraints():
    """force_first_iter: the first step's derived quantities are evaluated,
    without a history row (next_time is the first output time)."""
    _, its = _run(0.01, 0.1, 0.2)
    assert its[0]["constraints"] and not its[0]["row"]
    assert not any(it["constraints"] for it in its[1:8])
