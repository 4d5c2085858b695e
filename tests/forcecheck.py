"""Force comparison against the FP64 oracle with the cutoff band set apart.

The device decides r < r_c in FP32 on the fixed-point separation, FP64 decides
it exactly, so a pair within about 1e-7 (relative) of the cutoff can fall on
either side.  With a truncation whose force jumps at r_c (shifted, sharp) that
pair adds its full |F(r_c)| to one result and not to the other.  Such pairs are
a few in a million particles; the particles that hold one are compared with
the bound tol * rms + (their band pairs' |F|) instead of tol * rms.
"""
import numpy as np

BAND = 1e-6               # relative band of r^2 around r_c^2 (ten times the FP32 rounding)
TERM_TOL = 1e-6           # FP32 error relative to the sum of a particle's pair-force magnitudes


def force_errors(f, fref, pos, species, L, pot, oracle, tol):
    """(max_i |dF_i| / rms, median_i |dF_i| / |F_i|, band pairs) over the
    particles outside the band; asserts the band particles' bound.

    The max is the error relative to the rms force, the figure quoted for a
    result.  A particle whose pair forces are far larger than the rms (two
    particles pressed close) carries FP32 rounding in proportion to those
    terms: the bound asserted per particle is tol * rms + TERM_TOL * sum_j
    |F_ij|, and the max itself is returned for the caller to state."""
    rms = np.sqrt(np.mean(np.sum(fref ** 2, axis=1)))
    npair, jump = oracle.cutoff_band(pos, species, L, pot, BAND)
    band = jump > 0
    assert band.sum() <= max(8, 1e-3 * len(pos)), band.sum()
    err = np.max(np.abs(f - fref), axis=1)
    assert np.all(err[band] <= tol * rms + 1.001 * jump[band]), (err[band] - jump[band]).max() / rms
    scale = oracle.pair_force_scale(pos, species, L, pot)
    bound = tol * rms + TERM_TOL * scale
    assert np.all(err[~band] <= bound[~band]), ((err - bound)[~band].max(), rms)
    mx = err[~band].max() / rms
    norm = np.linalg.norm(fref, axis=1)
    ok = (norm > 1e-3 * rms) & ~band
    med = np.median(np.linalg.norm(f - fref, axis=1)[ok] / norm[ok])
    return mx, med, npair
