import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X GPU and the built libyee.so')


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope='session', autouse=True)
def tune_cache(tmp_path_factory):
    """Launch shapes tuned during the tests go to a cache file of the session, not the user's."""
    path = str(tmp_path_factory.mktemp('tune') / 'autotune.json')
    previous = os.environ.get('OPENSOL_AMD_TUNE_CACHE')
    os.environ['OPENSOL_AMD_TUNE_CACHE'] = path
    yield path
    if previous is None:
        os.environ.pop('OPENSOL_AMD_TUNE_CACHE', None)
    else:
        os.environ['OPENSOL_AMD_TUNE_CACHE'] = previous


@pytest.fixture(scope='session')
def gpu():
    """Skip-free guard: a gpu-marked test on a machine without a GPU fails loudly instead of passing silently."""
    if not gpu_available():
        pytest.fail('gpu test run without a visible GPU')
    from opensol_amd import _lib
    return _lib.load()
