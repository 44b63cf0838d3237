"""Training windows of the learned warm start (host logic, no GPU)."""
import pytest

from psb_project_amd.learned import window_starts


def test_windows_spread_over_the_horizon():
    w = window_starts(16, 200, 2000)
    assert w[0] == 0 and w[-1] + 200 == 2000
    assert len(w) == 16 and len(set(w)) == 16
    assert all(b - a == 120 for a, b in zip(w, w[1:]))


@pytest.mark.parametrize("count,steps,horizon", [(4, 200, None), (4, 200, 0), (4, 200, 200),
                                                 (4, 200, 150), (1, 200, 2000), (0, 200, 2000)])
def test_windows_from_the_start_without_a_longer_horizon(count, steps, horizon):
    assert window_starts(count, steps, horizon) == [0] * count


def test_windows_uneven_span_stays_inside():
    w = window_starts(7, 100, 1000)
    assert w[0] == 0 and w[-1] == 900 and w == sorted(w)


def test_record_refuses_a_warmup_list_of_the_wrong_length():
    from psb_project_amd import learned, scenes
    sc = scenes.pcisph_dam_break(n_side=8)
    with pytest.raises(ValueError, match="warm-up counts"):
        learned.record(None, [sc, sc], 3, None, warmup=[1])
