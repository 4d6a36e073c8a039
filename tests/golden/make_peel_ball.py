"""Copy the peel scene's initial particle positions out of the reference's
own run fixtures into ``peel_ball.npz``.

The reference keeps four peel runs (``anisompm_py/out/first_run`` and
``out/figs_spread``, ``run_correct.npz`` and ``run_wrong.npz``).  Each holds
``x_start``, the output of ``ball_particles`` (``src/peel_test.py:128-137``)
for the runs' configuration -- centre and spacing as recorded in the file
(0.5 and 1/36), R = 0.28 (the ``--R`` default), jitter 0.3 and seed 0 (the
defaults) -- before any step.  The four are identical; this copies
one, with its centre and spacing, as the builder's golden vector.  Only data
is read (``allow_pickle=False``); nothing of the reference runs.

Run:  python tests/golden/make_peel_ball.py
"""
from __future__ import annotations

import os

import numpy as np

SRC = "/root/reference/anisompm_py/out"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    runs = [os.path.join(SRC, d, f) for d in ("first_run", "figs_spread") for f in ("run_correct.npz", "run_wrong.npz")]
    xs = []
    for path in runs:
        with np.load(path, allow_pickle=False) as z:
            xs.append((z["x_start"].copy(), z["center"].copy(), float(z["spacing"])))
    x0, c0, s0 = xs[0]
    for x, c, s in xs[1:]:
        assert np.array_equal(x, x0) and np.array_equal(c, c0) and s == s0
    np.savez_compressed(os.path.join(HERE, "peel_ball.npz"), x_start=x0, center=c0, spacing=np.float64(s0),
                        R=np.float64(0.28), jitter=np.float64(0.3), seed=np.int64(0))
    print("peel_ball.npz:", x0.shape, c0, s0)


if __name__ == "__main__":
    main()
