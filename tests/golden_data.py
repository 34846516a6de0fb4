"""Loaders for the committed golden fixtures (written by tests/golden/make_golden.py)."""
import json
from functools import lru_cache
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / 'golden'


@lru_cache(None)
def load(name):
    return json.loads((GOLDEN / name).read_text())


def stream_runs():
    return load('stream_cells.json')


def complete_cases():
    return load('complete_only.json')


def firsthit_runs():
    return load('firsthit_runs.json')
