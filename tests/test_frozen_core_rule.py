"""eT's frozen-core orbital count per atom (ao_tool_class.F90:2425-2507), on the CPU."""
import pytest

from et_amd import ccsd
from et_amd.molecule import frozen_core_orbitals


def test_counts_follow_et_table():
    counts = {1: 0, 2: 0, 3: 0, 4: 0, 5: 1, 8: 1, 9: 1, 12: 1, 13: 5, 30: 5, 31: 9, 38: 9, 39: 14, 48: 14,
              49: 18, 70: 18, 71: 23, 80: 23, 81: 34, 103: 34, 104: 50, 105: 50}
    assert {z: frozen_core_orbitals(z) for z in counts} == counts
    assert sum(frozen_core_orbitals(z) for z in (1, 8, 9, 2)) == 2     # HOF-He: eT freezes 2


def test_beyond_table_is_an_error():
    with pytest.raises(ccsd.ETGError):
        frozen_core_orbitals(106)
