"""Bubble-configuration files (lattice_bubbles.rs:468-542) and the reference's example
configuration docs/examples/inputs/confY.txt (48 bubbles, periodic box L = 2), kept as
the fixture tests/golden/confY.txt."""
import os

import numpy as np
import pytest

from gw_fopt_amd import many_bubbles as mb

HERE = os.path.dirname(os.path.abspath(__file__))
CONF_Y = os.path.join(HERE, "golden", "confY.txt")


def conf_y_bubbles():
    lat = mb.CartesianLattice(origin=[0.0, 0.0, 0.0], basis=[[2.0, 0, 0], [0, 2.0, 0], [0, 0, 2.0]])
    lb = mb.LatticeBubbles(bubbles_interior=np.loadtxt(CONF_Y), lattice=lat)
    lb.with_boundary_condition("periodic")
    return lb


def test_conf_y_builds_like_the_reference_example():
    """docs/examples/generalized_bulkflow.py:741-752: CartesianLattice L = 2, periodic."""
    lb = conf_y_bubbles()
    assert lb.bubbles_interior.shape == (48, 4) and lb.bubbles_exterior.shape == (288, 4)
    assert lb.lattice_type == "Cartesian" and lb.volume == 8.0


@pytest.mark.parametrize("has_headers", [False, True])
def test_csv_round_trip(tmp_path, has_headers):
    lb = conf_y_bubbles()
    pi, pe = tmp_path / "int.csv", tmp_path / "ext.csv"
    lb.save_to_csv_files(pi, pe, has_headers=has_headers)
    text = pi.read_text().splitlines()
    assert (text[0] == "t,x,y,z") == has_headers
    back = mb.LatticeBubbles.from_csv_files(pi, pe, lb.lattice, has_headers=has_headers)
    assert back.bubbles_interior.tobytes() == lb.bubbles_interior.tobytes()
    assert back.bubbles_exterior.tobytes() == lb.bubbles_exterior.tobytes()


def test_csv_number_format(tmp_path):
    p = tmp_path / "b.csv"
    vals = np.array([[0.0, 1e-7, -2.5e300, 0.1]])
    mb.LatticeBubbles._write_csv(vals, p, False)
    assert p.read_text() == "0.0,1e-7,-2.5e300,0.1\n"
    np.testing.assert_array_equal(mb.LatticeBubbles.load_bubbles_from_csv(p), vals)


def test_csv_errors(tmp_path):
    p = tmp_path / "three.csv"
    p.write_text("0,1,2\n0,1,2\n")
    with pytest.raises(ValueError, match="Invalid number of columns"):
        mb.LatticeBubbles.load_bubbles_from_csv(p)
    q = tmp_path / "ragged.csv"
    q.write_text("0,1,2,3\n0,1,2\n")
    with pytest.raises(ValueError, match="deserialize_array2 error"):
        mb.LatticeBubbles.load_bubbles_from_csv(q)
    r = tmp_path / "tabs.csv"  # confY itself is tab separated: not a CSV bubble file
    r.write_text(open(CONF_Y).read())
    with pytest.raises(ValueError, match="Invalid number of columns"):
        mb.LatticeBubbles.load_bubbles_from_csv(r)
    e = tmp_path / "empty.csv"
    e.write_text("t,x,y,z\n")
    assert mb.LatticeBubbles.load_bubbles_from_csv(e, has_headers=True).shape == (0, 4)
