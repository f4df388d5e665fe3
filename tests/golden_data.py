"""Loads the golden fixtures of tests/golden/ (written by tests/golden/make_fixtures.py)."""
import json
import os

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name + ".json")) as f:
        return json.load(f)
