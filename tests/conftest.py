"""Test configuration: the `gpu` marker and the repository root on sys.path.

Tests marked `gpu` run the HIP path on an MI355X; the rest run on any CPU.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X GPU")
