"""Shared checks of engine results against the CPU oracle."""
import numpy as np

import pyoracle as O
import terranni_amd as T

REL_TOL = 1e-5  # heights vs the FP64 oracle, relative
# Denominator floor of the relative error (metres).  Where the stolen sites'
# integer heights cancel exactly the engine returns 0 while the FP64 mean of
# the oracle keeps ~1e-17 m of rounding noise.
REL_FLOOR = 1e-6


def rel_err(h, h64):
    """max and RMS relative error of the valid cells, and whether the
    NODATA sets agree."""
    valid = h64 != T.NODATA
    same_nodata = bool(np.array_equal(valid, h != T.NODATA))
    if not valid.any():
        return 0.0, 0.0, same_nodata
    e = np.abs(h[valid].astype(np.float64) - h64[valid]) / np.maximum(np.abs(h64[valid]), REL_FLOOR)
    return float(e.max()), float(np.sqrt(np.mean(e * e))), same_nodata


def check_tile(ctx, p, X, Y, Z, labels=True, s0=False):
    """Run one tile on the GPU and compare with the oracle; returns the
    oracle output for further checks."""
    h, lab, d2 = ctx.tile_run(p, X, Y, Z, labels=labels)
    ref = O.run(p, X, Y, Z, s0=s0)
    if labels:
        nbad = int((lab != ref["labels"]).sum())
        assert nbad == 0, f"{nbad} labels differ"
        assert np.array_equal(d2, ref["d2"])
    if p.accum_mode == T.ACC_S0:
        assert np.array_equal(h.view(np.uint32), ref["h_s0"].view(np.uint32))
    else:
        # the engine formula is exact: bit-equal to the oracle's h_fixed
        assert np.array_equal(h.view(np.uint32), ref["h_fixed"].view(np.uint32))
        emax, _, same = rel_err(h, ref["h64"])
        assert same
        assert emax <= REL_TOL, emax
    return h, ref
