"""CPU pins of the oracle's Hebbian rule (Hebbian::UpdateWeights,
src_wholebrain/hebbian.cu:3-69) by hand-derived scenarios on a single
synapse 0 -> 1: the window's end, an early target spike, a restart by a
second source spike, the same-call rules and the exp(-|w-1|) step size."""
import math

import numpy as np
import pytest

import oracle

OFF = np.array([0, 1, 1], np.int32)   # one synapse, 0 -> 1
TGT = np.array([1], np.int32)


def run(fires, window, inc=0.5, w0=1.0):
    """fires: list of sets of firing neurons per call; returns w and timer after each call."""
    timers = np.zeros(1, np.int32)
    w = np.array([w0], np.float32)
    out = []
    for f in fires:
        flags = np.zeros(2, np.uint8)
        flags[list(f)] = 1
        oracle.hebbian_update(OFF, TGT, flags, window, inc, timers, w)
        out.append((float(w[0]), int(timers[0])))
    return out


def close(a, b):
    return abs(a - b) <= 2 * np.spacing(np.float32(abs(b)))


def test_window_ends_without_target_spike():
    # W = 3: armed at call 0 (countdown 2 after the call), 1, then 0 at call 2: down by inc*exp(0)
    r = run([{0}, set(), set(), set()], 3)
    assert [x[1] for x in r] == [2, 1, 0, 0]
    assert r[0][0] == r[1][0] == 1.0 and r[2][0] == 0.5 and r[3][0] == 0.5


def test_target_spike_inside_window_goes_up():
    r = run([{0}, {1}, set(), set()], 3, w0=0.5)
    assert r[0] == (0.5, 2)
    assert close(r[1][0], 0.5 + 0.5 * math.exp(-0.5)) and r[1][1] == 0
    assert r[2] == r[1] and r[3] == r[1]  # stopped: no later change


def test_target_spike_at_the_window_end_is_too_late():
    # the countdown check comes first: at its last call the weight goes down
    r = run([{0}, set(), {1}], 3, w0=1.0)
    assert r[2] == (0.5, 0)


def test_second_source_spike_restarts_the_window():
    r = run([{0}, {0}, set(), set(), set()], 3)
    assert [x[1] for x in r] == [2, 2, 1, 0, 0]
    assert r[2][0] == 1.0 and r[3][0] == 0.5
    # a source spike on the last call of the window cancels the decrease
    r = run([{0}, set(), {0}, set(), set()], 3)
    assert [x[0] for x in r] == [1.0, 1.0, 1.0, 1.0, 0.5]


def test_same_call_rules():
    r = run([{0, 1}], 3, w0=2.0)  # both fire: up at once
    assert close(r[0][0], 2.0 + 0.5 * math.exp(-1.0)) and r[0][1] == 0
    r = run([{0, 1}], 1, w0=2.0)  # window 1: the countdown ends in the call: down, even if the target fires
    assert close(r[0][0], 2.0 - 0.5 * math.exp(-1.0)) and r[0][1] == 0


@pytest.mark.parametrize("w0", [0.0, 0.3, 1.0, 1.7, 3.0])
def test_step_size_is_inc_exp_minus_abs_w_minus_1(w0):
    r = run([{0}, set(), set()], 3, inc=0.25, w0=w0)
    exp = np.float32(w0) - np.float32(0.25 * math.exp(-abs(w0 - 1.0)))
    assert abs(r[2][0] - float(exp)) <= 2 * np.spacing(np.float32(max(abs(exp), 1e-3)))
