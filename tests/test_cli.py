"""Command line: hull CSV validation and option handling (host side)."""
import pytest

from tensordem_amd.cli import parser, read_hull_profile


def test_hull_profile_reader(tmp_path):
    good = tmp_path / "hull.csv"
    good.write_text("x,y\n-1.0,0.5\n0.0,0.0\n1.0,0.5\n")
    prof = read_hull_profile(good)
    assert prof.shape == (3, 2) and float(prof[2, 0]) == 1.0
    for text, msg in (("a,b\n1,2\n3,4\n", "x,y column"), ("x,y\n1,2\n", "at least two"),
                      ("x,y\n0,0\n0,0\n", "zero-length"), ("x,y\n0,nan\n1,1\n", "finite"),
                      ("x,y\n0,zz\n1,1\n", "line 2")):
        bad = tmp_path / "bad.csv"
        bad.write_text(text)
        with pytest.raises(ValueError, match=msg):
            read_hull_profile(bad)
    with pytest.raises(ValueError, match="cannot read"):
        read_hull_profile(tmp_path / "missing.csv")


def test_dem_2d_options_mirror_the_reference():
    a = parser().parse_args(["dem-2d"])
    assert (a.nx, a.ny, a.steps, a.dt, a.speed, a.save_every) == (21, 5, 2000, None, 1.0, 20)
    assert (a.breaking_strain, a.shear_breaking_strain) == (0.015, 0.03)
    a = parser().parse_args(["dem-2d", "--nx", "9", "--precision", "fp32",
                             "--dt-policy", "degree", "--steps", "10"])
    assert (a.nx, a.precision, a.dt_policy, a.steps) == (9, "fp32", "degree", 10)


def test_dem_2d_rejects_bad_counts_before_any_device_work(tmp_path):
    a = parser().parse_args(["dem-2d", "--steps", "0", "--output", str(tmp_path)])
    with pytest.raises(SystemExit):
        a.func(a)
