"""samos_nlist_pairs_noalloc (the allocation-free variant of samos_nlist_pairs)
against samos_nlist_pairs, through the C-ABI: bit-exact sorted pair rows on two
calls in a row, the second on a smaller system, then a list beyond the fixed
scratch maximum, which must be refused with an error and leave later calls
correct.
"""
import numpy as np
import pytest

from tests import scenarios as sc

pytestmark = pytest.mark.gpu


def system(gpu, st, rcut=2.4, pad=0.3):
    s = sc.load(gpu.System, st)
    s.nlist(rcut, pad)
    return s


def test_noalloc_pairs_equal_allocating_export(gpu):
    big = system(gpu, sc.plane_disks(6000, phi=0.8, seed=91))
    small = system(gpu, sc.sphere_disks(700, phi=0.9, seed=92))
    for s in (big, small, big):
        want = s.pairs()
        assert len(want) > 0
        got = s.pairs_noalloc(len(want) + 10)
        assert np.array_equal(got, want)
    # a capacity one row short of the count is an error, not a truncation
    with pytest.raises(gpu.SamosError):
        small.pairs_noalloc(len(small.pairs()) - 1)
    big.close()
    small.close()


def test_noalloc_pairs_refuses_beyond_its_maximum(gpu):
    # number density 4/pi and list radius 3.2: about 20 half pairs per
    # particle, ~3.1 M in all, beyond the 2,097,152 rows of the scratch array
    st = sc.plane_disks(150000, phi=4.0, seed=93)
    s = system(gpu, st, rcut=2.8, pad=0.4)
    with pytest.raises(gpu.SamosError, match="scratch maximum"):
        s.pairs_noalloc(1)
    s.close()
    after = system(gpu, sc.plane_disks(3000, phi=0.8, seed=94))
    assert np.array_equal(after.pairs_noalloc(100000), after.pairs())
    after.close()
