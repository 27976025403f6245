"""The oracle's pins catch plausible mistakes.

Each case builds the oracle with one planted mistake (``-DRO_MUTANT=k``, a
compile-time switch in oracle/raptr_oracle.c; the default build has none)
and runs the pins that must catch it: every listed pin has to fail on the
mistaken build.  A pin that passed on a wrong oracle would pin nothing.
"""
import os
import subprocess

import pytest

import test_oracle as pins

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

MUTANTS = {
    1: ("sign of the rotation term in the pixel trafo",
        ["test_trafo_matches_reference_build", "test_trafo_closed_form"]),
    2: ("id1z and id1y swapped in the channel index split",
        ["test_channel_index_round_trip", "test_workqueue_matches_reference_golden"]),
    3: ("pixel term dropped from the cull radius",
        ["test_workqueue_matches_reference_golden", "test_workqueue_is_conservative"]),
    4: ("|ray| factor dropped from the slab length",
        ["test_intersection_length_closed_forms", "test_intersection_length_brute_force"]),
    5: ("csr2ecsr gathers csr[i] instead of csr[yRowId[i]]",
        ["test_csr2ecsr_known_answer"]),
    6: ("normalisation x /= sum(x) dropped from the MLEM loop",
        ["test_mlem_fixed_point_and_normalisation"]),
    7: ("Philox with nine rounds", ["test_philox_known_answers"]),
    8: ("COO sort keyed (col, row) instead of (row, col)", ["test_coo_sort_reference_cases"]),
    9: ("point-line distance divided by |ap| instead of |ab|", ["test_distance_textbook"]),
}


def _build(k, out_dir):
    src = os.path.join(ORACLE_DIR, "raptr_oracle.c")
    out = os.path.join(out_dir, f"libraptr_oracle_m{k}.so")
    cmd = ["gcc", "-O1", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
           f"-DRO_MUTANT={k}", "-o", out, src, "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_pins_catch_planted_mistake(oracle, tmp_path, k):
    what, names = MUTANTS[k]
    path = _build(k, str(tmp_path))
    with oracle.swap_library(path):
        for name in names:
            with pytest.raises(AssertionError):
                getattr(pins, name)(oracle)
    # the unmodified build still passes the same pins
    for name in names:
        getattr(pins, name)(oracle)
