"""The oracle's hysteresis ramp (src/IrradianceProbes.cpp:462-476) pinned by what the mathematics
fixes: the closed form of the ramp, the one-step upload lag, the snap to the target, and when the
ramp moves for a given number of probes per update."""
import numpy as np

from oracle import py as orc


def test_ramp_closed_form_and_upload_lag():
    """Every probe updated every frame (count = P): the ramp moves before each update from the
    second on (h_j = T (1 - 0.9^j) after j moves, exact arithmetic), and the UBO receives the value
    before the move, so update k >= 1 blends with h_(k-1), not h_k, and update 0 with 0."""
    P, T = 64, 0.98
    used, _ = orc.hysteresis_ramp(P, [P] * 60, target=T)
    assert used[0] == 0.0 and used[1] == 0.0, "the first two updates blend with the initial 0"
    assert used[2] == float(np.float32(0.1) * np.float32(T)), "the first moved value reaches update 2, not 1"
    k = np.arange(1, 60)
    closed = T * (1 - 0.9 ** (k - 1))
    ramp = T - closed > 0.05  # values the ramp branch uploads (before the snap)
    got = np.array(used[1:])
    assert np.allclose(got[ramp], closed[ramp], atol=2e-6)
    # the snap: once |h - T| <= 0.05 the target is uploaded immediately and kept
    first_t = used.index(float(np.float32(T)))
    assert all(u == float(np.float32(T)) for u in used[first_t:])
    before = used[first_t - 1]
    assert T - before > 0.05 * 0.9 and T - before < 0.06, before
    assert first_t == 1 + int(np.ceil(np.log(0.05 / T) / np.log(0.9))), first_t


def test_ramp_moves_once_per_probe_count():
    """With c probes per update and P probes, the ramp has moved floor(c k / P) times when update k
    is recorded (the counter keeps the remainder when it subtracts P)."""
    for P, c in ((8, 3), (100, 37), (16384, 4096), (10, 10)):
        used, _ = orc.hysteresis_ramp(P, [c] * 80, target=0.98)
        moves = [0] + [int(used[k] != used[k - 1]) for k in range(1, 80)]
        # the upload lags one move: the value used changes at the move after the one that set it
        n_moves = np.cumsum(moves)
        expect = np.maximum(np.floor(c * np.arange(80) / P).astype(int) - 1, 0)
        # before the snap every move changes the uploaded value by >= 0.005
        ramp_phase = np.array(used) < 0.98 - 0.05
        assert np.array_equal(n_moves[ramp_phase], expect[ramp_phase]), (P, c)


def test_ramp_restarts_from_state():
    """Splitting a run in two and resuming from the returned (h, uploaded, counter) gives the same
    sequence (the function-static counter of the reference is per process; here it is state)."""
    counts = [5, 9, 3, 12, 7, 7, 1, 20, 4] * 12
    whole, _ = orc.hysteresis_ramp(16, counts)
    first, (h, up, upd) = orc.hysteresis_ramp(16, counts[:20])
    second, _ = orc.hysteresis_ramp(16, counts[20:], hysteresis=h, uploaded=up, updated=upd)
    assert whole == first + second
    # a fixed target is reached and held; the values used never exceed it and never decrease
    assert whole[-1] == float(np.float32(0.98))
    assert all(b >= a for a, b in zip(whole, whole[1:]))
