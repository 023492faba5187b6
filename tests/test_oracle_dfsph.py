"""A golden vector of the reference pins the oracle's shared f64 code.

The reference locks the FNV-1a hash of its canonical DFSPH tank after 40 steps
(docs/DETERMINISM.md:644-661, asserted by its CI, .github/workflows/ci.yml:596,
and by tests/test_fluid_rust.py:26). The oracle's DFSPH restatement runs on the
same functions the WCSPH oracle uses: mass_calibrated_to_rest, fill_box,
fill_box_boundary, recomputed_akinci_psi, the sorted uniform grid and its
candidate order, the cubic spline w and grad_w, boundary_sees_fluid, the Vec3
operation order and particles_hash64. Reproducing the hash bit for bit pins all
of them against the reference itself.
"""
import pyoracle as O

LOCKED_DFSPH_TANK_HASH = 0xF1F2825B5EEAAB0D


def test_canonical_dfsph_tank_hash_matches_the_reference():
    assert O.dfsph_tank_hash(40) == LOCKED_DFSPH_TANK_HASH


def test_the_hash_moves_with_the_trajectory():
    # a different step count gives a different hash: the check is not vacuous
    assert O.dfsph_tank_hash(39) != LOCKED_DFSPH_TANK_HASH
    assert O.dfsph_tank_hash(0) != O.dfsph_tank_hash(1)
