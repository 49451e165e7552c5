"""GPU test of sfp_extrapolate_velocity_noalloc, the entry point that allocates
nothing (its mask and sweep buffer live in a fixed device array): through the
C ABI it must give, bit for bit, what sfp_extrapolate_velocity_device gives
(Extrapolation::Solve as Simulation::ProjectVelocity calls it,
core/Simulation.cpp:236-240, core/Extrapolation.cpp:4-30), on two calls in a
row, the second on a smaller grid (the array keeps the first call's
contents), and free device memory must not change across a call."""
import numpy as np
import pytest

import sfp
from sfp import scenes

pytestmark = pytest.mark.gpu


def _inputs(n, seed, dev):
    import torch
    sc = scenes.random_pressure_scene(n, seed=seed)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    return sc, [t(v) for v in sc["velocity"]], t(sc["phi"]), [t(f) for f in sc["fraction"]]


@pytest.mark.parametrize("steps", [6, 3])
def test_noalloc_extrapolation_matches_the_context_entry_point(steps):
    import torch
    dev = torch.device("cuda:0")
    ptr = lambda ts: [x.data_ptr() for x in ts]  # noqa: E731
    stream = torch.cuda.current_stream().cuda_stream
    for call, n in enumerate([(40, 36, 44), (24, 20, 28)]):
        print(f"[case] no-alloc extrapolation call {call} grid {n} steps {steps}", flush=True)
        sc, vel, phi, frac = _inputs(n, seed=3 + call, dev=dev)
        ref = [v.clone() for v in vel]
        got = [v.clone() for v in vel]
        torch.cuda.synchronize()
        with sfp.Context.for_scene(sc) as ctx:
            ctx.extrapolate_velocity_device(ptr(ref), phi.data_ptr(), ptr(frac), steps)
        torch.cuda.synchronize()
        free_before = torch.cuda.mem_get_info()[0]
        sfp.extrapolate_velocity_noalloc(n, ptr(got), phi.data_ptr(), ptr(frac), steps, stream)
        torch.cuda.synchronize()
        if call > 0:  # the first call may load the library's device code
            assert torch.cuda.mem_get_info()[0] == free_before
        for a in range(3):
            assert torch.equal(got[a], ref[a]), (call, a)
        # the extrapolation changed something (faces next to liquid exist)
        assert any(not torch.equal(got[a], vel[a]) for a in range(3))


def test_noalloc_extrapolation_on_the_default_stream_and_its_limit():
    import torch
    dev = torch.device("cuda:0")
    ptr = lambda ts: [x.data_ptr() for x in ts]  # noqa: E731
    n = (16, 16, 16)
    sc, vel, phi, frac = _inputs(n, seed=9, dev=dev)
    ref = [v.clone() for v in vel]
    torch.cuda.synchronize()
    with sfp.Context.for_scene(sc) as ctx:
        ctx.extrapolate_velocity_device(ptr(ref), phi.data_ptr(), ptr(frac), 6)
    sfp.extrapolate_velocity_noalloc(n, ptr(vel), phi.data_ptr(), ptr(frac), 6, None)
    torch.cuda.synchronize()
    for a in range(3):
        assert torch.equal(vel[a], ref[a])
    # a grid beyond the fixed scratch is rejected before anything runs
    with pytest.raises(sfp.SfpError) as e:
        sfp.extrapolate_velocity_noalloc((400, 400, 400), ptr(vel), phi.data_ptr(), ptr(frac), 6)
    assert e.value.status == sfp.ERR_ARGUMENT
