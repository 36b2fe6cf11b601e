"""Launch-shape autotuner: times candidate tile shapes of an instance on the device and keeps the fastest.

Tile shapes never change results (every shape gives the same bits: halo copies round as their owners do), so the
tuner is free to pick any. For a 2D model it tries columns per lane x rows per tile; for a 3D model, planes per
tile. Each candidate runs on a throwaway engine over the same model (its own buffers), so the instance's state is
untouched; the best options are cached per (dimensions, shape, polarization) for the process.
"""
_CACHE = {}

CANDIDATES_2D = [(2, 0), (2, 16), (2, 32), (2, 64), (4, 0), (4, 32)]
CANDIDATES_3D = [0, 4, 8, 16, 32, 64]


def _key(model):
    return (len(model.shape), tuple(model.shape), getattr(model, 'polarization', '3d'))


def _time_engine(engine, steps, torch):
    engine.step(2)   # warm-up (code objects, tables)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    engine.step(steps)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def tune(sim, steps=10, candidates=None, use_cache=True):
    """Choose the launch options of a not-yet-run Simulation / Simulation3D by timing `steps` steps per candidate;
    sets them on the simulation and returns (options, {candidate: ms per step})."""
    import torch
    if sim.engine is not None:
        raise ValueError('tune() chooses launch options before the first run()')
    model = sim.model
    key = _key(model)
    if use_cache and key in _CACHE:
        sim._engine_options.update(_CACHE[key][0])
        return _CACHE[key]
    model.finalize()
    three = len(model.shape) == 3
    timings = {}
    if three:
        from .fdtd3d import Engine3D
        for planes in candidates or CANDIDATES_3D:
            engine = Engine3D(model, planes_per_tile=planes, capacity=steps + 4)
            try:
                timings[planes] = _time_engine(engine, steps, torch)
            finally:
                engine.close()
                del engine
        best = min(timings, key=timings.get)
        options = {'planes_per_tile': best}
    else:
        from .fdtd2d import Engine2D
        for lanes, rows in candidates or CANDIDATES_2D:
            engine = Engine2D(model, lane_columns=lanes, rows_per_tile=rows, capacity=steps + 4)
            try:
                timings[(lanes, rows)] = _time_engine(engine, steps, torch)
            finally:
                engine.close()
                del engine
        lanes, rows = min(timings, key=timings.get)
        options = {'lane_columns': lanes, 'rows_per_tile': rows}
    torch.cuda.empty_cache()
    sim._engine_options.update(options)
    _CACHE[key] = (options, timings)
    return options, timings


def clear_cache():
    _CACHE.clear()


__all__ = ['tune', 'clear_cache', 'CANDIDATES_2D', 'CANDIDATES_3D']
