"""Statistical checks of the Andersen heat bath on an ideal gas.

Restates test/unit/mdsim/integrators/verlet_nvt_andersen.cpp:101-174: the
centre-of-mass velocity and the temperature of the final state, their means
over the run, and the specific heat from the temperature fluctuations, each
within 4.5 standard deviations of the canonical value (a single check passes
with 99.999% probability).  Used for the oracle (CPU) and the device (GPU).

A velocity-rescaling thermostat passes the mean-temperature check but not
the specific heat (it suppresses the fluctuations), and a collision
probability other than rate * dt fails the collision count.
"""
import numpy as np

# the reference's set-up (verlet_nvt_andersen.cpp:177-200)
DENSITY = 0.3
TIMESTEP = 0.01
TEMPERATURE = 1.0
RATE = 10.0
BOX_RATIOS = (1.0, 2.0, 1.01)


def box_edges(npart):
    det = np.prod(BOX_RATIOS)
    edge = (npart / DENSITY / det) ** (1.0 / 3)
    return edge * np.asarray(BOX_RATIOS)


def sample_period():
    # independent samples: 3 collision times apart (verlet_nvt_andersen.cpp:108)
    return int(round(3.0 / (RATE * TIMESTEP)))


def check(temps, vcms, final_temp, final_vcm, npart, collisions=None, particle_steps=None, temp=TEMPERATURE):
    """temps: temperature per sample; vcms: (C, 3) centre-of-mass velocity per
    sample; final_*: of the last state; temp: of the heat bath.  Returns a
    dict of (value, bound)."""
    temps = np.asarray(temps, dtype=np.float64)
    vcms = np.asarray(vcms, dtype=np.float64)
    C = len(temps)
    out = {}
    # final state: v_cm (mean of N velocity components) and T (one variance)
    out["final_vcm"] = (np.max(np.abs(final_vcm)), 4.5 * np.sqrt(temp / (npart - 1)))
    # (relative bounds: <dT^2> / T^2 = 2 / (3 N); the reference divides once
    # more by T, which is the same at its T = 1)
    out["final_temp"] = (abs(final_temp - temp) / temp, 4.5 * np.sqrt(2.0 / (3 * npart)))
    # means over the run (N x C measurements of each velocity component)
    tol = 4.5 * np.sqrt(temp / (npart * C - 1))
    for k in range(3):
        m = vcms[:, k].mean()
        err = np.sqrt(vcms[:, k].var() / (C - 1))   # accumulator error_of_mean
        out[f"mean_vcm_{k}"] = (abs(m), tol)
        out[f"error_vcm_{k}"] = (err, tol)
    out["mean_temp"] = (abs(temps.mean() - temp) / temp, 4.5 * np.sqrt(2.0 / (3 * npart * (C - 1))))
    # specific heat per particle from the temperature fluctuations,
    # c_V = (3 N / 2)^2 <dT^2> / T^2 / N (k = 1)
    cv = (1.5 ** 2) * npart * temps.var() / temp ** 2
    cv_var = 1.5 * (3 + 6.0 / npart) / C
    out["c_V"] = (abs(cv - 1.5) / 1.5, 4.5 * np.sqrt(cv_var) / 1.5)
    if collisions is not None:
        # binomial count of heat-bath collisions, probability rate * dt each
        p = float(np.float32(RATE * TIMESTEP))
        mean = particle_steps * p
        out["collisions"] = (abs(collisions - mean), 4.5 * np.sqrt(particle_steps * p * (1 - p)))
    for name, (v, bound) in out.items():
        print(f"andersen {name}: {v:.4g} (bound {bound:.4g})")
    bad = {k: v for k, v in out.items() if not v[0] <= v[1]}
    assert not bad, bad
    return out
