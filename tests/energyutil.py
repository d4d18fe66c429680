"""Final-energy acceptance of long (chaotic) layout runs, shared by the GPU tests."""
import numpy as np


def perturbed(init, p):
    """Start p: the fixture init (p = 0) or a 1-ulp perturbation of every coordinate (the band
    fixture's scheme)."""
    if p == 0:
        return init.copy()
    rng = np.random.default_rng(100 + p)
    return np.nextafter(init, init + rng.choice([-1.0, 1.0], size=init.shape))


def energy_bars(e_def, e64, e_band):
    """The 600-iteration acceptance shared by the algo-2/4/8 tests.  The trajectory is chaotic,
    so each final energy is a sample.  Per start p the default path and the FP64 path (bit-exact
    with the strict reference, so its energy is the reference's from that start) run from the
    same positions.  Bars, each at max(1 %, 3 standard errors):
      * the mean paired difference (E_default - E_fp64) / E_fp64 over the starts;
      * the FP64 mean against the mean of the reference's own runs in the band fixture
        (standard error of the difference of two sample means)."""
    d = (e_def - e64) / e64
    se_d = d.std(ddof=1) / np.sqrt(len(d))
    m64, mb = e64.mean(), e_band.mean()
    se_b = np.sqrt(e64.var(ddof=1) / len(e64) + e_band.var(ddof=1) / len(e_band)) / mb
    rel_b = (m64 - mb) / mb
    return d.mean(), se_d, rel_b, se_b
