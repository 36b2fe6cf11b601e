"""The autotuner's host side: size classes, cache keys and the per-device JSON cache (round trip, atomic rewrite,
a corrupt or unwritable file)."""
import json
import os

from opensol_amd.autotune import TuneCache, class_key, size_class


class _Model:
    def __init__(self, shape, polarization=None):
        self.shape = shape
        self.polarization = polarization


def test_size_class_rounds_each_extent_up_to_a_power_of_two():
    assert size_class((128, 129, 1)) == (7, 8, 0)
    assert size_class((1024, 1000)) == (10, 10)
    assert size_class((64, 65, 192)) == (6, 7, 8)
    assert class_key('batch', _Model((300, 700), 'ez')) == 'batch:ez:9x10'
    assert class_key('single', _Model((100, 100, 100))) == 'single:3d:7x7x7'


def test_cache_round_trip_and_atomic_rewrite(tmp_path):
    path = str(tmp_path / 'sub' / 'tune.json')
    cache = TuneCache(path)
    assert cache.get('gfx950:256', 'batch:ez:9x9:lanes2') is None
    cache.put('gfx950:256', 'batch:ez:9x9:lanes2', {'rows_per_tile': 16}, {8: 1.5, 16: 1.25})
    cache.put('gfx950:256', 'batch:3d:7x7x7', {'planes_per_tile': 4}, {2: 3.0, 4: 2.5})
    with open(path) as f:
        data = json.load(f)
    assert data['gfx950:256']['batch:ez:9x9:lanes2'] == {'options': {'rows_per_tile': 16},
                                                         'ms': {'8': 1.5, '16': 1.25}}
    assert not [name for name in os.listdir(tmp_path / 'sub') if name.endswith('.tmp')]
    again = TuneCache(path)   # a new process reads the file
    assert again.get('gfx950:256', 'batch:3d:7x7x7') == {'planes_per_tile': 4}
    assert again.get('gfx942:304', 'batch:3d:7x7x7') is None


def test_corrupt_or_unwritable_cache_keeps_working(tmp_path):
    bad = tmp_path / 'bad.json'
    bad.write_text('{not json')
    cache = TuneCache(str(bad))
    assert cache.get('d', 'k') is None
    cache.put('d', 'k', {'rows_per_tile': 8}, {})
    assert TuneCache(str(bad)).get('d', 'k') == {'rows_per_tile': 8}
    blocked = tmp_path / 'file'
    blocked.write_text('')
    unwritable = TuneCache(str(blocked / 'tune.json'))   # its directory is a file
    unwritable.put('d', 'k', {'planes_per_tile': 2}, {})
    assert unwritable.get('d', 'k') == {'planes_per_tile': 2}
