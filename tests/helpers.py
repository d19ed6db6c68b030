"""Shared helpers for the parity tests (oracle = CPU restatement, test-only)."""

import numpy as np


def rel_fro(a, b):
    """||a - b||_F / ||b||_F."""
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
