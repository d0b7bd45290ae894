"""pm_widom_noalloc (the Widom entry point that allocates no device memory)
against pm_widom through the C ABI.

Both run the same kernels on the same image list, so the block sums must be
identical bit for bit; two no-alloc calls in a row, the second smaller than
the first, check that the fixed scratch array's leftover contents from one
call cannot leak into the next.  A run whose buffers exceed the array, and a
framework whose image list is not built yet, are refused on the host before
any device work."""
import math

import numpy as np
import pytest

from _parity import say

pytestmark = pytest.mark.gpu


def _handles(pm, case, probe_name):
    x, ff = case.crystal, case.ljff
    probe = case.probes[probe_name]
    xr = pm.replicate(x, pm.replication_factors(x.box, ff))
    fw = pm.device.framework(xr, ff)
    ew = None
    if probe.n_charges:
        ep = pm.setup_Ewald_sum(xr.box, math.sqrt(ff.r2_cutoff), eps=case.ewald_precision, verbose=False)
        ew = pm.device.ewald(xr, fw, ep)
    return fw, pm.device.forcefield(ff), ew, pm.device.probe(probe, ff)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_widom_noalloc_equals_widom(pm, precision):
    from porousmaterials_jl_amd._lib import PMError
    case = pm.synthetic.co2_widom_case()
    fw, ff, ew, pr = _handles(pm, case, "CO2")
    # first call on a fresh framework: its image list does not exist yet
    say(f"case: pm_widom_noalloc before the image list exists ({precision})")
    with pytest.raises(PMError, match="image list"):
        pm.device.widom_noalloc(fw, ff, ew, pr, 298.0, 5, 5, 100, precision)
    for seed, nb in ((21, 50000), (22, 3000)):
        say(f"case: pm_widom then pm_widom_noalloc, CO2 5 x {nb}, seed {seed} ({precision})")
        want = pm.device.widom(fw, ff, ew, pr, 298.0, seed, 5, nb, precision)
        got = pm.device.widom_noalloc(fw, ff, ew, pr, 298.0, seed, 5, nb, precision)
        assert np.array_equal(got, want), (seed, nb, got, want)
    # the same two no-alloc calls again, larger first, against the first results
    a = pm.device.widom_noalloc(fw, ff, ew, pr, 298.0, 21, 5, 50000, precision)
    b = pm.device.widom_noalloc(fw, ff, ew, pr, 298.0, 22, 5, 3000, precision)
    assert np.array_equal(a, pm.device.widom(fw, ff, ew, pr, 298.0, 21, 5, 50000, precision))
    assert np.array_equal(b, pm.device.widom(fw, ff, ew, pr, 298.0, 22, 5, 3000, precision))
    say("case: pm_widom_noalloc beyond its fixed scratch (refused on the host)")
    with pytest.raises(PMError, match="PM_NOALLOC_SCRATCH_BYTES"):
        pm.device.widom_noalloc(fw, ff, ew, pr, 298.0, 23, 5, 200000, precision)


def test_widom_noalloc_single_site_and_uncharged(pm):
    """A methane run (one LJ site, no Ewald handle) in the LJ grid framework,
    both precisions, two calls of falling size."""
    case = pm.synthetic.lj_grid_case()
    fw, ff, ew, pr = _handles(pm, case, "CH4")
    assert ew is None
    for precision in ("mixed", "fp64"):
        for seed, nb in ((31, 40000), (32, 500)):
            say(f"case: CH4 5 x {nb}, seed {seed}: pm_widom vs pm_widom_noalloc ({precision})")
            want = pm.device.widom(fw, ff, None, pr, 298.0, seed, 5, nb, precision)
            got = pm.device.widom_noalloc(fw, ff, None, pr, 298.0, seed, 5, nb, precision)
            assert np.array_equal(got, want), (precision, seed, nb)
