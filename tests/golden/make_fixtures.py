"""Writes the fixtures of this folder from the reference's own data files.

control_function_ohara.npy: the reference's default drive
(assets/control_function_ohara.npy, float32, 2001 samples at 2 kHz, read by
ControlFunction::from_model_config, model/functional/control.rs:174-223).
It is loaded with numpy's pickle-free loader and saved unchanged.

Run from the repository root: python3 tests/golden/make_fixtures.py
"""
import os

import numpy as np

REFERENCE = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    raw = np.load(os.path.join(REFERENCE, "assets", "control_function_ohara.npy"),
                  allow_pickle=False)
    assert raw.dtype == np.float32 and raw.ndim == 1, (raw.dtype, raw.shape)
    np.save(os.path.join(HERE, "control_function_ohara.npy"), raw, allow_pickle=False)
    print("control_function_ohara.npy", raw.shape, raw.dtype)


if __name__ == "__main__":
    main()
