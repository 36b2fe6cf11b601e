"""yee2d_batch_step_static, the no-allocation batch entry point, called through the C ABI, against the batch context
(yee2d_batch_create + yee_step over a caller buffer): the same fields and flux spectra, bit for bit, on two calls in a
row, the second with fewer and smaller instances (the static array keeps the first call's descriptors past the
second call's count); more instances than the array holds are refused."""
import ctypes

import numpy as np
import pytest

from tests.test_gpu_batch import INSTANCES, _make, _snapshot

pytestmark = pytest.mark.gpu

SECOND = [((20, 26), 4, 'hz', None, 77), ((31, 17), 5, 'ez', 'periodic_y', 45), ((19, 19), 4, 'ez', None, 90)]


def _static_run(lib, batch, counts):
    """Step batch.simulations by counts[k] steps with yee2d_batch_step_static (descriptors as Batch2D builds them)."""
    from opensol_amd import _lib
    torch = batch.torch
    descs = (_lib.Desc2DC * len(batch.simulations))()
    for k, (sim, count) in enumerate(zip(batch.simulations, counts)):
        engine = sim.engine
        engine._ensure_capacity(engine.n_step + count)
        descs[k] = batch._cached_descriptor(engine)
        descs[k].step_limit = engine.n_step + count
    stream = torch.cuda.current_stream(batch.simulations[0].engine.device).cuda_stream
    _lib.check(lib.yee2d_batch_step_static(descs, len(descs), max(counts), stream), 'yee2d_batch_step_static')
    for sim, count in zip(batch.simulations, counts):
        sim.engine.n_step += count


def _compare(specs, seed0):
    from opensol_amd import _lib
    from opensol_amd.batch import Batch2D
    lib = _lib.load()
    counts = [spec[4] for spec in specs]
    plain = [_make(spec, seed0 + k) for k, spec in enumerate(specs)]
    static = [_make(spec, seed0 + k) for k, spec in enumerate(specs)]
    plain_batch = Batch2D([sim for sim, _ in plain])
    static_batch = Batch2D([sim for sim, _ in static])   # builds the descriptors; its own context is never made
    print(f'no-alloc: {len(specs)} instances, steps {counts}', flush=True)
    plain_batch.run(counts)
    _static_run(lib, static_batch, counts)
    for (a, ma), (b, mb), count in zip(plain, static, counts):
        fa, fluxa = _snapshot(a, ma)
        fb, fluxb = _snapshot(b, mb)
        assert a.engine.n_step == b.engine.n_step == count
        assert any(np.any(values != 0) for values in fa.values())
        for name in fa:
            np.testing.assert_array_equal(fb[name], fa[name])
        np.testing.assert_array_equal(fluxb, fluxa)
    plain_batch.close()
    for sim, _ in plain + static:
        sim.close()


def test_static_batch_equals_batch_context_on_two_calls(gpu):
    _compare(INSTANCES, 0)
    _compare(SECOND, 50)


def test_static_batch_refuses_more_instances_than_it_holds(gpu):
    from opensol_amd import _lib
    lib = _lib.load()
    descs = (_lib.Desc2DC * (_lib.STATIC_BATCH_MAX + 1))()
    status = lib.yee2d_batch_step_static(descs, len(descs), 1, None)
    assert status != 0
    assert b'at most 256 instances' in lib.yee_last_error()
