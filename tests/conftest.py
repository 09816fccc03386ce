import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950)")


@pytest.fixture(scope="session")
def wiki_vote_path():
    return os.path.join(ROOT, "tests", "golden", "wiki-vote.g")
