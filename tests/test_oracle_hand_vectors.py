"""Hand-computed vectors for the bookkeeping rules of the path (oracle; the same scenes on the
device in tests/test_gpu_hand_vectors.py).

* Adaptive sweep, BoundDispatcher::processBody (pkg/common/Dispatching.cpp:30-62):
  sweepLength = max(minSweepDistFactor*sweepDist, min(max(0.9*sweepLength, newLength), sweepDist)),
  newLength = dist*targetInterv/(iter - lastUpdateIter), dist = |pos - refPos|_inf, 0 when
  dist == 0; the first call sees refPos = NaN (core/Bound.hpp:26).  The collider runs on the
  step after NewtonIntegrator saw |pos - refPos|_inf >= sweepLength (NewtonIntegrator.cpp:70-83,
  InsertionSortCollider.cpp:177-189).  Velocities and the step are powers of two, so every
  displacement below is exact and the expected numbers are worked out by hand in the comments.
* Swap-with-last erase, InteractionContainer::erase and conditionalyEraseNonReal
  (core/InteractionContainer.cpp:86-91, .hpp:104-112): the hole takes the last interaction,
  which is tested again before the scan moves on.

The bodies carry no walls and are thin along x, so that boxes (isotropic, r_max) overlap
while the bodies never touch.  Such an interaction becomes real (the Ig2 functor creates its
geometry, Ip2 its physics) with zero depth; the law's strict box test requests its erasure
once the boxes part (Superquadrics.cpp:1194-1207), and the next collider run erases it.
"""
import numpy as np
import pytest

from sudodem_amd import scene as S

DT = 2.0 ** -10
VD = 2.0 ** -8  # verletDist


def _thin_body(mat, x, vx):
    b = S.NewSuperquadrics_rot2(2.0 ** -12, 2.0 ** -10, 2.0 ** -10, 1.0, 1.0, mat, 1, 0, 0, 0)
    b.state.pos = np.array([x, 0.0, 0.0])
    b.state.vel = np.array([vx, 0.0, 0.0])
    return b


def _engines(verlet):
    return [S.ForceResetter(), S.InsertionSortCollider([S.Bo1_Superquadrics_Aabb(), S.Bo1_Wall_Aabb()],
                                                       verletDist=verlet),
            S.InteractionLoop([S.Ig2_Wall_Superquadrics_SuperquadricsGeom(),
                               S.Ig2_Superquadrics_Superquadrics_SuperquadricsGeom()],
                              [S.Ip2_SuperquadricsMat_SuperquadricsMat_SuperquadricsPhys()], [S.SuperquadricsLaw()]),
            S.NewtonIntegrator(damping=0.0, gravity=(0.0, 0.0, 0.0), isSuperquadrics=1)]


def _omega(lib, xs, vs, verlet=VD, dt=DT):
    mat = S.SuperquadricsMat(Kn=1e5, Ks=7e4, frictionAngle=0.5, density=2650)
    O = S.Omega(lib=lib)
    O.materials.append(mat)
    O.bodies.extend([_thin_body(mat, x, v) for x, v in zip(xs, vs)])
    O.engines = _engines(verlet)
    O.dt = dt
    return O


def sweep_history(lib, slow_down_at=None, n=60):
    """Three bodies 1 m apart (never in one box): A at 2^-4 m/s, B at 2^-6 m/s, C at rest.
    Runs n steps (with `slow_down_at`, A's velocity becomes 2^-7 m/s after that step) and
    returns per step (iter after the step, collider runs, sweepLength, refPos x,
    lastUpdateIter) of the three bodies."""
    O = _omega(lib, [0.5, 1.5, 2.5], [2.0 ** -4, 2.0 ** -6, 0.0])
    O._open()
    rows = []
    for k in range(1, n + 1):
        O.run(1)
        rows.append((int(O.get("scene.iter")[0]), O.counters()["collider_runs"], O.get("bound.sweepLength").copy(),
                     O.get("bound.refPos").reshape(-1, 3)[:, 0].copy(), O.get("bound.lastUpdateIter").copy()))
        if slow_down_at is not None and k == slow_down_at:
            v = O.get("state.vel").reshape(-1, 3)
            v[0, 0] = 2.0 ** -7
            O.set("state.vel", v)
    return rows


A_STEP, B_STEP = 2.0 ** -14, 2.0 ** -16  # displacement per step (2^-4 and 2^-6 m/s times 2^-10 s)


def check_first_runs(rows):
    """The first three steps all run the collider (rows[k] is the state after step k + 1, i.e.
    after iteration k): at iter 0 and 1 strideActive is still false (it is set only once
    NewtonIntegrator.maxVelocitySq, NaN until Newton's first step, is >= 0,
    InsertionSortCollider.cpp:279, NewtonIntegrator.hpp:75), at iter 2 fastestBodyMaxDist is
    still at its initial -1 (InsertionSortCollider.cpp:181, .hpp:219).
      run 1 (iter 0) sees refPos = NaN: dist = NaN, newLength = NaN, max(0.9 * 0, NaN) = 0
        (std::max keeps its first argument), so every body gets 0.1 * verletDist;
      run 2 (iter 1), one move later: A: dist = 2^-14, newLength = 2^-14 * 50 / (1 - 0) =
        50 * 2^-14 = 3.0517578125e-3 (above 0.9 * s0 and 0.1 * verletDist, below verletDist =
        3.90625e-3); B: 50 * 2^-16 = 7.62939453125e-4; C: dist = 0 -> sweepLength = 0;
      run 3 (iter 2): one move over one iteration: the same lengths."""
    s0 = 0.1 * VD
    it, runs, sl, ref, lui = rows[0]
    assert (it, runs) == (1, 1)
    assert sl.tolist() == [s0, s0, s0]
    assert ref.tolist() == [0.5, 1.5, 2.5] and lui.tolist() == [0, 0, 0]
    for k in (1, 2):
        it, runs, sl, ref, lui = rows[k]
        assert (it, runs) == (k + 1, k + 1)
        assert sl.tolist() == [50 * A_STEP, 50 * B_STEP, 0.0]
        assert ref.tolist() == [0.5 + k * A_STEP, 1.5 + k * B_STEP, 2.5]
        assert lui.tolist() == [k, k, k]


def test_adaptive_sweep_constant_velocity(oracle):
    rows = sweep_history(oracle)
    check_first_runs(rows)
    # A and B reach 50 moves beyond refPos (set after 2 moves) at Newton of iter 51: the
    # collider runs at iter 52, dist = 50 moves over 52 - 2 iterations: lengths unchanged
    for it, runs, *_ in rows[3:52]:
        assert runs == 3, it
    it, runs, sl, ref, lui = rows[52]
    assert (it, runs) == (53, 4)
    assert sl.tolist() == [50 * A_STEP, 50 * B_STEP, 0.0]
    assert ref.tolist() == [0.5 + 52 * A_STEP, 1.5 + 52 * B_STEP, 2.5]
    assert lui.tolist() == [52, 52, 52]


def test_adaptive_sweep_decay_and_cap(oracle):
    """A slows to 2^-7 m/s (2^-17 per step) after its second move.  B still triggers the run
    at iter 52; A's dist there = 50 * 2^-17 over 50 iterations, newLength = 50 * 2^-17 =
    3.8e-4 < 0.9 * 50 * 2^-14 = 2.7e-3: A's length decays by the factor 0.9 only.  A fast body
    (2^-2 m/s, 2^-12 per step) gets newLength = 2^-12 * 50 / 1 = 0.0122 at run 2, above
    verletDist: the cap."""
    rows = sweep_history(oracle, slow_down_at=2)
    check_first_runs(rows)
    for it, runs, *_ in rows[3:52]:
        assert runs == 3, it
    it, runs, sl, ref, lui = rows[52]
    assert (it, runs) == (53, 4)
    assert sl.tolist() == [0.9 * (50 * A_STEP), 50 * B_STEP, 0.0]
    assert ref.tolist() == [0.5 + 2 * A_STEP + 50 * 2.0 ** -17, 1.5 + 52 * B_STEP, 2.5]
    assert lui.tolist() == [52, 52, 52]
    O = _omega(oracle, [0.5], [2.0 ** -2])
    O._open()
    O.run(2)
    assert O.counters()["collider_runs"] == 2
    assert O.get("bound.sweepLength").tolist() == [VD]
    assert O.get("bound.lastUpdateIter").tolist() == [1]


def erase_sequence(lib, vel):
    """Five thin bodies on the x axis 1.5 mm apart (boxes of half-width 1 mm overlap their
    neighbours only).  The first collider run (sort on x, then the sweep) creates
    (0,1), (1,2), (2,3), (3,4) in this order.  Returns the container after the first step and
    after 80 more steps with the given x velocities (0.8 mm at 1 m/s and dt = 1e-5 s)."""
    O = _omega(lib, [0.0, 1.5e-3, 3.0e-3, 4.5e-3, 6.0e-3], vel, verlet=1e-6, dt=1e-5)
    O._open()
    O.run(1)
    i1, i2, real = O.interactions()
    first = list(zip(i1.tolist(), i2.tolist()))
    O.run(80)
    i1, i2, real = O.interactions()
    return first, list(zip(i1.tolist(), i2.tolist()))


# A body moving at 1 m/s parts from a resting neighbour's box after 0.5 mm = 50 steps (plus
# the sweep enlargement, the same for every moving body): in each case the pairs named below
# part at the same step, are requested for erasure by the law there and erased by one scan of
# the next collider run (the collider runs every step: verletDist 1e-6 < 1e-5 per step).
ERASE_CASES = [
    # (0,1) and (2,3) separate at the same run: the scan erases (0,1) at position 0 and moves
    # the last, (3,4), into it; (3,4) stays, (1,2) stays, (2,3) is now last: removed.
    ([-1.0, 0.0, -1.0, 0.0, 0.0], [(0, 1), (1, 2), (2, 3), (3, 4)], [(3, 4), (1, 2)]),
    # (0,1) and (3,4) at the same run: (3,4) moves into position 0 and is tested again there:
    # erased too, and (2,3) takes its place.
    ([-1.0, 0.0, 0.0, 0.0, 1.0], [(0, 1), (1, 2), (2, 3), (3, 4)], [(2, 3), (1, 2)]),
    # (1,2) alone (bodies 0 and 1 move left together): (3,4) moves into position 1.
    ([-1.0, -1.0, 0.0, 0.0, 0.0], [(0, 1), (1, 2), (2, 3), (3, 4)], [(0, 1), (3, 4), (2, 3)]),
]


@pytest.mark.parametrize("vel,first,after", ERASE_CASES)
def test_swap_with_last_erase(oracle, vel, first, after):
    got_first, got_after = erase_sequence(oracle, vel)
    assert got_first == first
    assert got_after == after
