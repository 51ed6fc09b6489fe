"""Trajectory frames downloaded asynchronously in id order (md_frames_*, the data of the reference's
writeXyzFrame, Trajectory.cpp:4-25): each frame row k is bit-equal to the position and velocity of the particle
whose id is k (the inverse of the cell re-sort permutation, Trajectory.cpp:9-11); a capture holds the state as it
stood at the call even while later steps run; slots are reused only once released; ids outside 0..n-1 and states
larger than the frames are rejected. The XYZ text itself is checked byte for byte against the reference's writer
in the facade suite (TrajectoryAsync tests) and against the reference's md_sim in tests/test_gpu_md_sim.py.
"""
import numpy as np
import pytest

import mdgpu

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)


def by_id(state):
    """The state's positions and velocities as an (n, 6) array in id order, from plain downloads."""
    pos = state.download(mdgpu.FIELD_POS).reshape(-1, 4)[:, :3]
    vel = state.download(mdgpu.FIELD_VEL).reshape(3, -1).T
    ids = state.download(mdgpu.FIELD_ID).astype(np.int64)
    out = np.empty((state.n, 6))
    out[ids, :3] = pos
    out[ids, 3:] = vel
    return out


def liquid(ctx, cells=8, steps=200):
    s = mdgpu.fcc_system(ctx, cells, 0.8442, 0.728, 11)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(s)
    nl.forces(s, LJ, mdgpu.MD_FP64)
    nl.run(s, LJ, 0.005, steps, precision=mdgpu.MD_FP64)
    return s, nl


def test_frame_rows_are_the_particles_in_id_order(ctx):
    s, _ = liquid(ctx)
    ids = s.download(mdgpu.FIELD_ID)
    assert not np.array_equal(ids, np.arange(s.n))  # re-sorted by the list builds
    f = mdgpu.Frames(ctx, s.n, slots=2)
    slot = f.capture(s)
    frame, box = f.wait(slot)
    assert box == s.box and frame.shape == (s.n, 6)
    assert np.array_equal(frame, by_id(s))
    f.release(slot)


def test_capture_holds_the_state_while_later_steps_run(ctx):
    s, nl = liquid(ctx)
    f = mdgpu.Frames(ctx, s.n, slots=3)
    expected, slots = [], []
    for _ in range(3):
        expected.append(by_id(s))
        slots.append(f.capture(s))
        # enqueued after the capture; the copy of the frame overlaps these steps
        nl.run(s, LJ, 0.005, 100, precision=mdgpu.MD_MIXED, record=False)
    for want, slot in zip(expected, slots):
        got, _ = f.wait(slot)
        assert np.array_equal(got, want)
        f.release(slot)
    assert not np.array_equal(expected[0], expected[2])


def test_slots_are_reused_only_after_release(ctx):
    s, _ = liquid(ctx, cells=5, steps=20)
    f = mdgpu.Frames(ctx, s.n, slots=2)
    a = f.capture(s)
    b = f.capture(s)
    assert (a, b) == (0, 1)
    with pytest.raises(mdgpu.MdError):
        f.capture(s)
    f.release(a)  # released without a wait: the copy is waited for inside
    assert f.capture(s) == 0
    for slot in (b, 0):
        frame, _ = f.wait(slot)
        assert np.array_equal(frame, by_id(s))
        f.release(slot)
    with pytest.raises(mdgpu.MdError):
        f.wait(1)  # released: holds no frame


def test_bad_ids_and_oversized_states_are_rejected(ctx):
    s, _ = liquid(ctx, cells=5, steps=20)
    small = mdgpu.Frames(ctx, s.n - 1, slots=1)
    with pytest.raises(mdgpu.MdError):
        small.capture(s)
    ids = s.download(mdgpu.FIELD_ID).copy()
    ids[7] = s.n  # outside 0..n-1: the scatter skips it and flags the frame
    s.upload(mdgpu.FIELD_ID, ids)
    f = mdgpu.Frames(ctx, s.n, slots=1)
    slot = f.capture(s)
    with pytest.raises(mdgpu.MdError, match="permutation"):
        f.wait(slot)
    f.release(slot)


def test_frames_of_a_smaller_state_use_its_own_count(ctx):
    big, _ = liquid(ctx, cells=6, steps=20)
    s, _ = liquid(ctx, cells=5, steps=20)
    f = mdgpu.Frames(ctx, big.n, slots=1)
    for st in (big, s):
        slot = f.capture(st)
        frame, box = f.wait(slot)
        assert frame.shape == (st.n, 6) and box == st.box
        assert np.array_equal(frame, by_id(st))
        f.release(slot)
