"""Launch-shape autotuner: times candidate tile shapes on the device and keeps the fastest, per size class, in a JSON
cache per device.

Tile shapes never change results (every shape gives the same bits: halo copies round as their owners do), so the
tuner is free to pick any. The choices are made per size class: the rounded-up log2 of each grid extent, with the
polarisation for 2D (a 300 x 700 E_z grid is class ez:9x10), separately for a single instance and for a member of a
batched launch, whose kernels differ:

  * single 2D instance: columns per lane x rows per tile, timed on a throwaway engine over the same model;
  * single 3D instance: x planes per tile, likewise;
  * batch member, 2D: rows per tile (columns per lane are shared by the whole batch), timed as a one-member batch
    (Batch2D over a throwaway engine, the same kernels a real batch launches);
  * batch member, 3D: x planes per tile, timed as a one-member Batch3D.

The cache file is JSON, {device: {class key: {"options": ..., "ms": {candidate: ms per step}}}}, with the device named
by its architecture and compute-unit count (e.g. "gfx950:256"). Its path is OPENSOL_AMD_TUNE_CACHE, else
~/.cache/opensol_amd/autotune.json; it is written atomically after every new class. Batch2D / Batch3D take
autotune=True, FDTDServer(autotune=True) tunes every admitted job's class before its engine exists, and bench.py
tunes the batched workload's members before its timed region.

The reference has no counterpart: its launch shapes are fixed per kernel (physics/em/solvers/fdtd_2d_base.py:508-599).
"""
import json
import math
import os
import threading

CANDIDATES_2D = [(2, 0), (2, 16), (2, 32), (2, 64), (4, 0), (4, 32)]
CANDIDATES_3D = [0, 4, 8, 16, 32, 64]
BATCH_ROWS_2D = [8, 16, 32, 64]
BATCH_PLANES_3D = [2, 4, 8]


def size_class(shape):
    """Rounded-up log2 of each grid extent."""
    return tuple(int(math.ceil(math.log2(max(int(n), 1)))) for n in shape)


def class_key(kind, model):
    """Cache key of a model's size class for `kind` ('single' or 'batch')."""
    shape = tuple(model.shape)
    pol = '3d' if len(shape) == 3 else model.polarization
    return f'{kind}:{pol}:' + 'x'.join(str(c) for c in size_class(shape))


def default_cache_path():
    return os.environ.get('OPENSOL_AMD_TUNE_CACHE') or os.path.join(os.path.expanduser('~'), '.cache', 'opensol_amd',
                                                                    'autotune.json')


class TuneCache:
    """The per-device JSON cache of tuned options (loaded on first use; thread-safe)."""

    def __init__(self, path=None):
        self.path = path or default_cache_path()
        self._data = None
        self._lock = threading.Lock()

    def _load(self):
        if self._data is None:
            try:
                with open(self.path) as f:
                    data = json.load(f)
                self._data = data if isinstance(data, dict) else {}
            except (OSError, ValueError):
                self._data = {}
        return self._data

    def get(self, device, key):
        with self._lock:
            entry = self._load().get(device, {}).get(key)
        return None if entry is None else dict(entry['options'])

    def put(self, device, key, options, timings):
        with self._lock:
            data = self._load()
            data.setdefault(device, {})[key] = {'options': dict(options),
                                                'ms': {str(c): float(t) for c, t in timings.items()}}
            try:   # an unwritable location keeps the choices for this process only
                directory = os.path.dirname(os.path.abspath(self.path))
                os.makedirs(directory, exist_ok=True)
                tmp = f'{self.path}.{os.getpid()}.tmp'
                with open(tmp, 'w') as f:
                    json.dump(data, f, indent=1, sort_keys=True)
                os.replace(tmp, self.path)
            except OSError:
                pass

    def clear(self):
        with self._lock:
            self._data = {}


_DEFAULT = None


def default_cache():
    global _DEFAULT
    if _DEFAULT is None or _DEFAULT.path != default_cache_path():
        _DEFAULT = TuneCache()
    return _DEFAULT


def device_key(torch, device=None):
    props = torch.cuda.get_device_properties(torch.cuda.current_device() if device is None else device)
    return f'{getattr(props, "gcnArchName", props.name).split(":")[0]}:{props.multi_processor_count}'


def _elapsed_ms(run, steps, torch):
    run(2)   # warm-up (code objects, tables)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    run(steps)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def _batch_ms(batch, steps, torch):
    """ms per step of `steps` batched steps (descriptor upload outside the timed launches)."""
    batch.run(2)   # warm-up
    batch.prepare(steps)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    batch.launch()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def _lookup(key, cache, use_cache, torch):
    cache = cache or default_cache()
    device = device_key(torch)
    return cache, device, (cache.get(device, key) if use_cache else None)


def tune(sim, steps=10, candidates=None, use_cache=True, cache=None):
    """Choose the launch options of a not-yet-run Simulation / Simulation3D by timing `steps` steps per candidate on
    a throwaway engine (or take its size class's cached choice); sets them on the simulation and returns
    (options, {candidate: ms per step}) ({} timings for a cached choice)."""
    import torch
    if sim.engine is not None:
        raise ValueError('tune() chooses launch options before the first run()')
    model = sim.model
    key = class_key('single', model)
    cache, device, options = _lookup(key, cache, use_cache, torch)
    if options is not None:
        sim._engine_options.update(options)
        return options, {}
    model.finalize()
    timings = {}
    if len(model.shape) == 3:
        from .fdtd3d import Engine3D
        for planes in candidates or CANDIDATES_3D:
            engine = Engine3D(model, planes_per_tile=planes, capacity=steps + 4)
            try:
                timings[planes] = _elapsed_ms(engine.step, steps, torch)
            finally:
                engine.close()
        options = {'planes_per_tile': min(timings, key=timings.get)}
    else:
        from .fdtd2d import Engine2D
        for lanes, rows in candidates or CANDIDATES_2D:
            engine = Engine2D(model, lane_columns=lanes, rows_per_tile=rows, capacity=steps + 4)
            try:
                timings[(lanes, rows)] = _elapsed_ms(engine.step, steps, torch)
            finally:
                engine.close()
        lanes, rows = min(timings, key=timings.get)
        options = {'lane_columns': lanes, 'rows_per_tile': rows}
    torch.cuda.empty_cache()
    sim._engine_options.update(options)
    cache.put(device, key, options, timings)
    return options, timings


class _Member:
    """A stand-in for a Simulation in a one-member timing batch: the model and a throwaway engine over it."""

    def __init__(self, model, engine):
        self.model = model
        self.engine = engine
        self._engine_options = {}

    def finalize(self):
        pass


def tune_batch_member(sim, steps=10, lane_columns=2, candidates=None, use_cache=True, cache=None):
    """Choose the options of a not-yet-run simulation as a member of a batched launch (2D: rows per tile at the
    batch's columns per lane; 3D: x planes per tile) by timing one-member batches, or take its class's cached choice;
    sets them on the simulation and returns (options, timings)."""
    import torch
    from .batch import Batch2D, Batch3D
    if sim.engine is not None:
        raise ValueError('tune_batch_member() chooses launch options before the engine exists')
    model = sim.model
    three = len(model.shape) == 3
    key = class_key('batch', model) + ('' if three else f':lanes{int(lane_columns)}')
    cache, device, options = _lookup(key, cache, use_cache, torch)
    if options is not None:
        sim._engine_options.update(options)
        return options, {}
    model.finalize()
    timings = {}
    if three:
        from .fdtd3d import Engine3D
        for planes in candidates or BATCH_PLANES_3D:
            engine = Engine3D(model, planes_per_tile=planes, capacity=steps + 8)
            batch = Batch3D([_Member(model, engine)])
            try:
                timings[planes] = _batch_ms(batch, steps, torch)
            finally:
                batch.close()
                engine.close()
        options = {'planes_per_tile': min(timings, key=timings.get)}
    else:
        from .fdtd2d import Engine2D
        for rows in candidates or BATCH_ROWS_2D:
            engine = Engine2D(model, lane_columns=lane_columns, rows_per_tile=rows, capacity=steps + 8)
            batch = Batch2D([_Member(model, engine)], lane_columns=lane_columns)
            try:
                timings[rows] = _batch_ms(batch, steps, torch)
            finally:
                batch.close()
                engine.close()
        options = {'rows_per_tile': min(timings, key=timings.get)}
    torch.cuda.empty_cache()
    sim._engine_options.update(options)
    cache.put(device, key, options, timings)
    return options, timings


def clear_cache():
    """Forget the default cache's in-memory entries (the file stays)."""
    default_cache().clear()


__all__ = ['tune', 'tune_batch_member', 'size_class', 'class_key', 'TuneCache', 'default_cache', 'clear_cache',
           'CANDIDATES_2D', 'CANDIDATES_3D', 'BATCH_ROWS_2D', 'BATCH_PLANES_3D']
