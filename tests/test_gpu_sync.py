"""Synchronization probe (mcp_time_sync): a grid-wide barrier inside a
cooperative persistent kernel against a dependent launch in a captured graph.

The probe itself checks after every barrier that each workgroup sees its
neighbour's writes (a stale read is an error) and bounds every wait (a wait
that runs out of spins is an error, not a hang); the test requires positive
times from it and a refusal for more workgroups than can be resident.
"""
import pytest

import mcpdlp as mp

pytestmark = pytest.mark.gpu


def test_grid_barrier_and_graph_launch_costs():
    e = mp.Engine(0)
    try:
        for g in (1, 8, 64):
            bu, lu = e.time_sync(g, 400)
            print(f"{g} workgroups: barrier {bu:.2f} us, graph launch {lu:.2f} us", flush=True)
            assert bu > 0.0 and lu > 0.0
        with pytest.raises(RuntimeError, match="resident"):
            e.time_sync(1 << 24, 4)
    finally:
        e.close()
