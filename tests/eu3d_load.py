"""Imports the package directory eu3d-gpu_amd (its name is not an identifier)."""
import importlib
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
eu3d = importlib.import_module("eu3d-gpu_amd")
compare_states = importlib.import_module("eu3d-gpu_amd.compare_states")
