"""Zero-copy torch views of the resident state (the reference's shared_buffer_to_dlpack,
python/bindings.cpp:282-340): the views alias the device buffers cn_step writes, follow later
steps, hold the context past close() and refuse a reallocating reserve()."""
import gc

import numpy as np
import pytest

import cnhip
from cnhip import scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def params():
    return cnhip.default_params(dt=1 / 120, substeps=4, default_iterations=4, max_fluid_neighbors=64)


def test_torch_views_alias_resident_state():
    s = cnhip.Solver(0)
    sc = scenes.dam_break_64k(side=10)
    s.reserve(sc.n, 64)
    s.upload(sc)
    s.step(params())
    v = s.torch_views()
    dv = s.device_views()
    pos, vel = v["positions_inv_mass"], v["velocities"]
    assert pos.is_cuda and pos.shape == (sc.n, 4) and pos.dtype == torch.float32
    assert pos.data_ptr() == dv.positions_inv_mass and vel.data_ptr() == dv.velocities
    torch.cuda.current_stream().wait_stream(s.torch_stream())
    p0, _, v0 = s.readback()
    np.testing.assert_array_equal(pos.cpu().numpy(), p0)
    np.testing.assert_array_equal(vel.cpu().numpy(), v0)
    s.step(params())  # the same tensors now hold the next state
    torch.cuda.current_stream().wait_stream(s.torch_stream())
    p1, _, v1 = s.readback()
    assert not np.array_equal(p0, p1)
    np.testing.assert_array_equal(pos.cpu().numpy(), p1)
    np.testing.assert_array_equal(vel.cpu().numpy(), v1)
    with pytest.raises(RuntimeError):
        s.reserve(sc.n * 4, 64)
    s.close()  # deferred: the views still alias the buffers
    assert s._ctx
    np.testing.assert_array_equal(pos.cpu().numpy(), p1)
    del v, pos, vel
    gc.collect()
    assert not s._ctx


def test_served_session_tensors_equal_readback():
    from cnhip import serving
    from test_serving import small_sessions
    ses = small_sessions()
    srv = serving.Server(ses, params(), budget_bytes=1 << 40, shards=2, graphs=True)
    try:
        done = srv.serve(serving.poisson_arrivals(300.0, 0.1, range(len(ses)), seed=3))
        assert done
        for k in range(len(ses)):
            (pos, vel), stream = srv.session_tensors(k)
            torch.cuda.current_stream().wait_stream(stream)
            p, _, v = srv.session_state(k)
            np.testing.assert_array_equal(pos.cpu().numpy(), p, err_msg=f"session {k}")
            np.testing.assert_array_equal(vel.cpu().numpy(), v, err_msg=f"session {k}")
    finally:
        srv.close()
