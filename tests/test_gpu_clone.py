"""GPU tests of plan clones (raptr_recon_clone): a second plan that shares
the matrix layouts and owns its work vectors.  A clone run on its own stream
at the same time as its parent gives, bit for bit, the images of the same
work run alone; kernel shapes are per plan; the parent outlives its
clones."""
import numpy as np
import pytest

from test_gpu_recon import small_system  # noqa: F401

pytestmark = pytest.mark.gpu


def _side_by_side(R, C, Ya, Yb, x0, niter, F, use_graph):
    """R on one stream with counts Ya and C on another with Yb, at once;
    returns both images."""
    from raptr_amd.device import DeviceArray, Event, Stream
    s1, s2 = Stream(), Stream()
    dya, dyb = DeviceArray.from_numpy(Ya.reshape(-1)), DeviceArray.from_numpy(Yb.reshape(-1))
    dxa, dxb = DeviceArray.from_numpy(x0.reshape(-1)), DeviceArray.from_numpy(x0.reshape(-1))
    s1.sync(), s2.sync()
    R.run(dya, dxa, niter, F, use_graph=use_graph, stream=s1.handle)
    C.run(dyb, dxb, niter, F, use_graph=use_graph, stream=s2.handle)
    done = Event()
    done.record(s2.handle)
    s1.wait(done)
    s1.sync()
    return dxa.numpy(), dxb.numpy()


@pytest.mark.parametrize("F", [1, 64])
def test_clone_runs_beside_its_parent(gpu, small_system, F):
    """One frame (the fused iteration) and 64 frames (grouped layouts on the
    matrix cores): parent and clone run different counts side by side on
    two streams, eagerly and from graphs; each image equals that plan's
    work run alone, bit for bit."""
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    kw = dict(row_order=S["setup"]) if F == 1 else dict(row_order=S["setup"], grid=S["grid"],
                                                         group=16)
    R = Reconstructor(S["A"], S["sens"], "f32", F, **kw)
    C = R.clone()
    assert R.clone_count() == 1
    rng = np.random.default_rng(3)
    y = S["y"].astype(np.float32)
    Ya = (y[:, None] * rng.uniform(0.5, 2.0, (1, F))).astype(np.float32)
    Yb = (y[:, None] * rng.uniform(0.5, 2.0, (1, F))).astype(np.float32)
    x0 = np.full((n, F), 1.0 / n, np.float32)
    if F == 1:
        Ya, Yb, x0 = Ya[:, 0].copy(), Yb[:, 0].copy(), x0[:, 0].copy()
    alone_a = R.mlem(Ya, x0, 5)
    alone_b = R.mlem(Yb, x0, 5)
    for graph in (False, True):
        a, b = _side_by_side(R, C, Ya, Yb, x0, 5, F, graph)
        np.testing.assert_array_equal(a.reshape(alone_a.shape), alone_a)
        np.testing.assert_array_equal(b.reshape(alone_b.shape), alone_b)
    C.close()
    assert R.clone_count() == 0


def test_clone_shapes_and_lifetime(gpu, small_system):
    """A clone starts with its parent's kernel shapes and changes them on
    its own; a clone of a clone counts against the first plan; the parent
    cannot be destroyed while clones live."""
    from raptr_amd import RaptrError
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    R = Reconstructor(S["A"], S["sens"], "f32", 1, row_order=S["setup"])
    R.set_fused(2, 8, 64, 32)
    C = R.clone()
    assert C.fused() == R.fused()
    C.set_fused(False)
    assert R.fused()["on"] == 2 and C.fused()["on"] == 0
    a, c = R.mlem(S["y"], x0, 5), C.mlem(S["y"], x0, 5)
    R.set_fused(False)
    np.testing.assert_array_equal(R.mlem(S["y"], x0, 5), c)  # the same separate passes
    assert float(np.linalg.norm(a - c) / np.linalg.norm(c)) <= 1e-6
    D = C.clone()
    assert R.clone_count() == 2
    with pytest.raises(RaptrError):
        R.close()
    D.close()
    C.close()
    assert R.clone_count() == 0
    R.close()
