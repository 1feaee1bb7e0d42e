"""The LM update loop's host logic on the CPU (disorder_amd/csrc/lm_host.hpp,
Reconstruction/LMsolver.m:126-215) with the device solve and trial energies
replaced by scripted stubs (tests/native/lm_host_harness.cpp, built here with
g++): finite scripts follow the reference's bookkeeping step by step, and
NaN/Inf energies or motion end the call instead of spinning."""
import ctypes
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LM_OK, LM_NONFINITE_T, LM_NONFINITE_ENERGY, LM_NO_CONVERGENCE = 0, 1, 2, 3


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("lmh") / "lmh.so")
    subprocess.run([gxx, "-std=c++17", "-O1", "-shared", "-fPIC", "-o", out,
                    os.path.join(HERE, "native", "lm_host_harness.cpp")], check=True)
    lib = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.harness_lm.restype = ctypes.c_int
    lib.harness_lm.argtypes = [ctypes.c_int, dp, dp, ctypes.POINTER(ctypes.c_uint8), dp, ctypes.c_double, dp, dp,
                               dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), dp]
    lib.harness_restrict.restype = ctypes.c_double
    lib.harness_restrict.argtypes = [ctypes.c_double]
    return lib


def run(lib, T, w, cT, EnPrev, steps, trial, tL=(0.05, 0.02), winit=1e-3):
    T = np.asfortranarray(np.array(T, np.float64))  # [NM x 6] column-major, as the C-ABI takes it
    NM = T.shape[0]
    w = np.array(w, np.float64)
    cT = np.array(cT, np.uint8)
    tl = np.array(tL, np.float64)
    EnPrev = np.array(EnPrev, np.float64)
    steps = np.ascontiguousarray(steps, np.float64)
    trial = np.ascontiguousarray(np.atleast_2d(trial), np.float64)
    rounds, ntr = ctypes.c_int(0), ctypes.c_int(0)
    En = np.zeros(NM)
    dp = ctypes.POINTER(ctypes.c_double)
    Tf = T.reshape(-1, order="F").copy()
    st = lib.harness_lm(NM, Tf.ctypes.data_as(dp), w.ctypes.data_as(dp),
                        cT.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), tl.ctypes.data_as(dp), winit,
                        EnPrev.ctypes.data_as(dp), steps.ctypes.data_as(dp), trial.ctypes.data_as(dp), trial.shape[0],
                        ctypes.byref(rounds), ctypes.byref(ntr), En.ctypes.data_as(dp))
    return st, Tf.reshape(6, NM).T, w, cT.astype(bool), rounds.value, ntr.value, En


def test_scripted_accept_reject_matches_reference_bookkeeping(harness):
    # two states: state 0 rejected in round 1 (energy up) and accepted in
    # round 2; state 1 accepted in round 1 and not re-evaluated
    s0 = np.array([0.4, -0.2, 0.1, 0.01, 0.0, -0.02])
    s1 = np.array([0.001, 0.0, 0.0, 0.0, 0.0, 0.0])
    steps = np.stack([s0, s1])
    T0 = np.zeros((2, 6))
    trial = [[2.0, 0.5], [0.9, 99.0]]  # row r: energies of round r
    st, T, w, cT, rounds, ntr, En = run(harness, T0, [1e-3, 1e-3], [0, 0], [1.0, 1.0], steps, trial)
    assert st == LM_OK and rounds == 2 and ntr == 3  # round 1: both states; round 2: state 0 only
    # state 0: w doubled once (LMsolver.m:200), then /1.2 on success (:202)
    assert w[0] == pytest.approx(2e-3 / 1.2, rel=1e-15)
    assert w[1] == pytest.approx(1e-3 / 1.2, rel=1e-15)
    # state 0 takes the step of its second solve (w = 2e-3: half the first),
    # state 1 keeps its latest step (solved again in round 2 with unchanged w)
    np.testing.assert_allclose(T[0], s0 * 0.5, rtol=1e-15)
    np.testing.assert_allclose(T[1], s1, rtol=1e-15)
    # convergence (:209): max |dt| < tL(1) and max |dtheta| (deg) < tL(2)
    assert not cT[0] and cT[1]
    np.testing.assert_allclose(En, [0.9, 0.5])


def test_converged_states_are_left_alone(harness):
    steps = np.ones((2, 6)) * 0.3
    st, T, w, cT, rounds, ntr, _ = run(harness, np.zeros((2, 6)), [1e-3, 5.0], [1, 0], [0.0, 1.0], steps,
                                       [[7.0, 0.5]])
    assert st == LM_OK and rounds == 1 and ntr == 1
    np.testing.assert_array_equal(T[0], 0.0)
    assert w[0] == 1e-3  # converged: no damping change (:202)
    assert w[1] == pytest.approx(5.0 / 1.2)


def _doublings_to_cutoff(w0):
    k, w = 0, w0
    while w <= 1e10:
        w *= 2
        k += 1
    return k


@pytest.mark.parametrize("bad", [math.nan, math.inf])
def test_nonfinite_trial_energy_is_a_rejection_and_the_loop_ends(harness, bad):
    # the reference would spin here: NaN compares false both ways
    steps = np.ones((1, 6)) * 0.01
    st, T, w, cT, rounds, ntr, En = run(harness, np.zeros((1, 6)), [1e-3], [0], [1.0], steps, [[bad]])
    k = _doublings_to_cutoff(1e-3)
    assert st == LM_OK
    assert rounds == k + 1  # k rejections, then w > 1e10 keeps EnPrev (:195) and accepts
    assert w[0] == pytest.approx(1e-3 * 2 ** k / 1.2)
    assert np.all(np.isfinite(T)) and np.max(np.abs(T)) < 1e-12  # the last step was damped to nothing
    assert En[0] == 1.0


def test_finite_rejections_end_at_the_damping_cutoff(harness):
    steps = np.ones((1, 6)) * 0.01
    st, _, w, _, rounds, _, _ = run(harness, np.zeros((1, 6)), [1e-8], [0], [1.0], steps, [[2.0]])
    assert st == LM_OK and rounds == _doublings_to_cutoff(1e-8) + 1 <= 62


def test_nonfinite_energy_at_current_motion_is_an_error(harness):
    steps = np.ones((2, 6)) * 0.01
    st, T, *_ = run(harness, np.zeros((2, 6)), [1e-3, 1e-3], [0, 0], [1.0, math.nan], steps, [[0.5, 0.5]])
    assert st == LM_NONFINITE_ENERGY
    np.testing.assert_array_equal(T, 0.0)  # nothing applied


@pytest.mark.parametrize("where", ["T", "w"])
def test_nonfinite_motion_or_damping_is_an_error(harness, where):
    T = np.zeros((2, 6))
    w = [1e-3, 1e-3]
    if where == "T":
        T[1, 4] = math.inf
    else:
        w[0] = math.nan
    st, *_ = run(harness, T, w, [0, 0], [1.0, 1.0], np.zeros((2, 6)), [[0.5, 0.5]])
    assert st == LM_NONFINITE_T


def test_singular_steps_take_no_step(harness):
    steps = np.full((1, 6), math.nan)
    st, T, w, cT, rounds, _, _ = run(harness, np.ones((1, 6)) * 0.1, [1e-3], [0], [1.0], steps, [[1.0]])
    assert st == LM_OK and rounds == 1
    np.testing.assert_array_equal(T, 0.1)
    assert cT[0]  # a zero step is below both limits


def test_restrict_rotation(harness):
    f = harness.harness_restrict
    assert f(3.5) == pytest.approx(3.5 - 2 * math.pi)
    assert f(-7.0) == pytest.approx(-7.0 + 2 * math.pi)
    assert abs(f(1e300)) <= math.pi  # bounded for far-out values
    assert math.isnan(f(math.nan)) and math.isinf(f(math.inf))  # returned as is, no loop
