"""The persistent-kernel group barrier on device atomics, stress-tested as in
the reference's tests/test_multiblock.cu:8-31 (independent groups, uneven
exits, a head start that deadlocks a grid-wide phase barrier, per-thread
publication checked every generation)."""
import sys

import pytest

import gpuboscia_amd as gb

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("groups,group_blocks,threads", [(64, 4, 256), (128, 2, 64), (256, 1, 128),
                                                         (8, 32, 256)])
def test_group_barrier_stress(groups, group_blocks, threads):
    say(f"[barrier] groups={groups} k={group_blocks} threads={threads}")
    errors, timeouts, per_sync = gb.grid_barrier_selftest(groups, group_blocks, threads,
                                                          phases_base=2000, phase_step=37)
    say(f"[barrier] errors={errors} timeouts={timeouts} {per_sync * 1e6:.2f} us/sync")
    assert timeouts == 0
    assert errors == 0


def test_bounded_wait_reports_timeout():
    # A spin limit of one poll makes waiters give up at once: the kernel must
    # still exit and report the timeout instead of hanging.
    errors, timeouts, _ = gb.grid_barrier_selftest(16, 4, 64, phases_base=200, phase_step=0,
                                                   spin_limit=1)
    say(f"[barrier] forced: errors={errors} timeouts={timeouts}")
    assert timeouts >= 1


def test_rejects_non_resident_grid():
    with pytest.raises(gb.InvalidArgument):
        gb.grid_barrier_selftest(1 << 16, 16, 1024, phases_base=200)
