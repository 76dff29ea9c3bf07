"""The Fortran MD driver's deck reader, on the CPU: tags it refuses stop the
run before any device work, as the reference's loaders stop it."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.golden.configs import mixture_deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "genesis6_amd", "lib", "g6md")


def run(deck):
    if not os.path.exists(EXE):
        pytest.skip("g6md is not built")
    with tempfile.TemporaryDirectory() as d:
        return subprocess.run([EXE, "fp64"], input=deck, capture_output=True, text=True,
                              timeout=60, cwd=d)


def small_deck():
    com = np.array([[0.0, 0.0, 0.0], [3.0, 0.0, 0.0]])
    q = np.array([[1.0, 0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0]])
    return mixture_deck(com, q, 18.0, 8.5, 0.5, 0.002, 1, 1)


def test_pcn2_is_refused_as_the_reference_refuses_it():
    """Andersen_ReadPCN2 (AndersenModule.F90:118-132) stops with a message
    that points to @PCN4."""
    deck = small_deck().replace("@ID08", "@PCN2\n1 1.0 1e-4 0.0\n@ID08", 1)
    r = run(deck)
    assert r.returncode != 0
    assert "@PCN2 is not supported any more. Please use @PCN4 instead." in r.stderr


def test_non_cubic_box3_reaches_the_library():
    """An orthorhombic @BOX3 (BoxModule.F90:109-116) is read and handed to
    g6_init_box3 with its three edges; without a device the library's
    error code stops the run."""
    deck = small_deck().replace("@BXLA\n18.0", "@BOX3\n18.0 18.0 19.0", 1)
    r = run(deck)
    assert r.returncode != 0 and "g6_init_box3 failed with code -4" in r.stderr


def second_component(deck, n=2, model="OPLSMEOH"):
    rows = "\n".join("%r %r 5.0 1.0 0.0 0.0 0.0" % (6.0 + 3.0 * k, 6.0) for k in range(n))
    return deck + "@ID08\n%s\n@NX4A\n%d\n%s\n" % (model, n, rows)


def test_mixture_deck_reaches_the_library():
    """Two @ID08/@NX4A blocks make a mixture (Mixture.F90:636-712): both
    components are read and the first is handed to g6_init_box3; without a
    device the library's error code stops the run."""
    r = run(second_component(small_deck()))
    assert r.returncode != 0 and "g6_init_box3 failed with code -4" in r.stderr


@pytest.mark.parametrize("extra", ["@PCN4\n1 1.0 1e-4\n", "@SNPO\n5\n"])
def test_mixture_refuses_barostat_and_snapshots(extra):
    """A mixture takes neither the barostat nor snapshots here: the run
    stops before any device work."""
    deck = small_deck().replace("@ID08", extra + "@ID08", 1)
    r = run(second_component(deck))
    assert r.returncode != 0 and "a mixture takes neither" in r.stderr


def test_third_component_and_ncmp_beyond_two_are_refused():
    r = run(second_component(second_component(small_deck())))
    assert r.returncode != 0 and "at most two components" in r.stderr
    r = run(small_deck().replace("@ID08", "@NCMP\n3\n@ID08", 1))
    assert r.returncode != 0 and "one or two components" in r.stderr
