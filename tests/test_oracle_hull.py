"""Pins of the oracle's polyline hull law (oracle/dem2d.c `hull_particle`,
restating hull_contact.py:7-75) against the reference's own known answers:

* tests/test_hull_contact.py:7-36 -- force (0, 10) N at 0.01 m overlap with
  k = 1000, action-reaction, endpoint contact, damping never attracts;
* docs/hull-contact-verification/report.json (tests/golden/
  hull_contact_verification.json) -- the sagitta error table over 10
  discretisations x 6 penetrations (max / mean relative magnitude error,
  dropout fraction, direction error), the corner contact reducing to
  circle-circle overlap, and the degenerate disk-centre-on-the-hull case
  whose canonical-tangent fallback normal (hull_contact.py:53-68) gives
  (-0.0, 50.0) N for either digitisation order;
* closed forms: the sagitta law |F - k d| / (k d) = s / d at chord
  midpoints (tests/test_hull_contact_verification.py:44-59), the
  approach-oriented fallback normal worked out by hand.

Each pin is also run against a planted mistake (a plausible misreading of
the law), which it must reject.
"""
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.plate import hull_forces

REPORT = json.loads((GOLDEN / "hull_contact_verification.json").read_text())
ZERO = np.zeros(2)


def bow(radius, segments, arc=2 * math.pi, phase=0.0):
    """Inscribed polygon on an arc (hull_contact_verification.py:50-64)."""
    ang = np.arange(segments + 1, dtype=np.float64) * (arc / segments) + phase
    return np.stack([radius * np.cos(ang), radius * np.sin(ang)], axis=1)


def sagitta(radius, segments):
    return radius * (1.0 - math.cos(math.pi / segments))


def test_known_force_and_action_reaction():
    p = np.array([[0.0, 0.04], [0.5, 0.20]])
    f, reaction = hull_forces(p, np.zeros_like(p), [[-1.0, 0.0], [1.0, 0.0]], ZERO,
                              particle_radius=0.05, hull_radius=0.0, stiffness=1000.0)
    assert f[0] == pytest.approx([0.0, 10.0], abs=1e-12)
    assert np.all(f[1] == 0.0)
    assert np.all(f.sum(axis=0) + reaction == 0.0)


def test_endpoint_contact_is_radial_from_the_vertex():
    f, _ = hull_forces([[1.03, 0.0]], [[0.0, 0.0]], [[0.0, 0.0], [1.0, 0.0]], ZERO,
                       particle_radius=0.05, hull_radius=0.0, stiffness=100.0)
    assert f[0, 0] == pytest.approx(100.0 * (0.05 - 0.03), rel=1e-12)
    assert f[0, 1] == 0.0


def test_damping_does_not_create_attraction():
    f, _ = hull_forces([[0.0, 0.04]], [[0.0, 10.0]], [[-1.0, 0.0], [1.0, 0.0]], ZERO,
                       particle_radius=0.05, hull_radius=0.0, stiffness=100.0, damping=100.0)
    assert np.all(f == 0.0)
    # approaching instead: k d + c |v_n|
    f, _ = hull_forces([[0.0, 0.04]], [[0.0, -0.5]], [[-1.0, 0.0], [1.0, 0.0]], ZERO,
                       particle_radius=0.05, hull_radius=0.0, stiffness=100.0, damping=100.0)
    assert f[0, 1] == pytest.approx(100.0 * 0.01 + 100.0 * 0.5, rel=1e-12)


def test_hull_radius_and_hull_velocity_enter_the_law():
    """Overlap r_p + r_h - d; normal speed relative to the moving hull."""
    f, _ = hull_forces([[0.0, 0.06]], [[0.0, 0.0]], [[-1.0, 0.0], [1.0, 0.0]], [0.0, 0.2],
                       particle_radius=0.05, hull_radius=0.02, stiffness=100.0, damping=10.0)
    assert f[0, 1] == pytest.approx(100.0 * 0.01 + 10.0 * 0.2, rel=1e-12)


def _scan_row(level, fraction, radius=0.5, rp=0.05, k=1000.0, per_segment=4):
    """One row of the reference's discretisation scan
    (hull_contact_verification.py:197-251) evaluated with the oracle."""
    verts = bow(radius, level)
    vang = np.arange(level, dtype=np.float64) * (2.0 * math.pi / level)
    blocks = [vang, vang + math.pi / level]
    for kk in range(max(per_segment - 2, 0)):
        blocks.append((np.arange(level, dtype=np.float64) + 0.25 * (kk + 1))
                      * (2.0 * math.pi / level))
    ang = np.concatenate(blocks)
    overlap = fraction * rp
    dist = rp + radius - overlap
    pos = np.stack([dist * np.cos(ang), dist * np.sin(ang)], axis=1)
    f, _ = hull_forces(pos, np.zeros_like(pos), verts, ZERO, particle_radius=rp,
                       hull_radius=0.0, stiffness=k)
    # exact circular hull: radial normal, magnitude k * overlap
    nrm = np.sqrt((pos * pos).sum(axis=1))
    exact = (k * np.maximum(rp + radius - nrm, 0.0))[:, None] * (pos / nrm[:, None])
    em = np.sqrt((exact * exact).sum(axis=1))
    mag = np.sqrt((f * f).sum(axis=1))
    rel = np.abs(mag - em) / em
    d = (np.arctan2(f[:, 1], f[:, 0]) - np.arctan2(exact[:, 1], exact[:, 0]) + math.pi) \
        % (2.0 * math.pi) - math.pi
    d = np.where(mag > 0, np.abs(d), math.pi)
    return dict(dropout_fraction=float((mag == 0).mean()),
                max_relative_magnitude_error=float(rel.max()),
                mean_relative_magnitude_error=float(rel.mean()),
                max_direction_error_deg=math.degrees(float(d.max())))


def test_discretisation_error_table_matches_reference_report():
    rows = REPORT["error_table"]
    assert len(rows) == 60
    for g in rows:
        got = _scan_row(g["segments"], g["penetration_fraction"])
        assert got["dropout_fraction"] == g["dropout_fraction"], g
        for key in ("max_relative_magnitude_error", "mean_relative_magnitude_error"):
            assert got[key] == pytest.approx(g[key], rel=1e-9, abs=1e-12), (key, g)
        assert got["max_direction_error_deg"] == pytest.approx(g["max_direction_error_deg"],
                                                               rel=1e-8, abs=1e-9), g


@pytest.mark.parametrize("segments", [16, 32, 64])
@pytest.mark.parametrize("fraction", [0.5, 0.1])
def test_sagitta_law_at_chord_midpoints(segments, fraction):
    radius, rp, k = 0.5, 0.05, 1000.0
    overlap = fraction * rp
    ang = np.arange(segments, dtype=np.float64) * (2 * math.pi / segments) + math.pi / segments
    dist = rp + radius - overlap
    pos = np.stack([dist * np.cos(ang), dist * np.sin(ang)], axis=1)
    f, _ = hull_forces(pos, np.zeros_like(pos), bow(radius, segments), ZERO,
                       particle_radius=rp, hull_radius=0.0, stiffness=k)
    measured = float(np.max(np.abs(np.sqrt((f * f).sum(axis=1)) - k * overlap) / (k * overlap)))
    predicted = sagitta(radius, segments) / overlap
    if predicted >= 1.0:
        assert measured == 1.0
    else:
        assert measured == pytest.approx(predicted, abs=1e-10)


def test_corner_contact_is_circle_to_circle():
    gold = REPORT["vertex_check"]
    verts = bow(0.5, 24, arc=math.pi)
    apex = verts[0]
    overlap = 0.2 * 0.05
    pos = (apex + apex / np.sqrt(apex @ apex) * (0.05 - overlap))[None, :]
    f, reaction = hull_forces(pos, np.zeros_like(pos), verts, ZERO, particle_radius=0.05,
                              hull_radius=0.0, stiffness=1000.0)
    mag = float(np.sqrt(f[0] @ f[0]))
    assert mag == pytest.approx(gold["analytic_force_N"], rel=1e-12)
    assert mag == pytest.approx(gold["kernel_force_N"], rel=1e-13)
    assert np.all(f.sum(axis=0) + reaction == 0.0)


def test_degenerate_centre_on_hull_uses_canonical_fallback():
    """hull_contact.py:53-68 and the reference report: a centre exactly on the
    line gets the normal of the canonical tangent (first nonzero component
    positive), oriented against the approach, for either digitisation order."""
    inv = REPORT["invariance"]
    kw = dict(particle_radius=0.05, hull_radius=0.0, stiffness=1000.0, damping=10.0)
    hull = np.array([[-1.0, 0.0], [1.0, 0.0]])
    on = np.zeros((1, 2))
    fwd, _ = hull_forces(on, on, hull, ZERO, **kw)
    bwd, _ = hull_forces(on, on, hull[::-1].copy(), ZERO, **kw)
    assert list(fwd[0]) == inv["degenerate_forward_force"]
    assert list(bwd[0]) == inv["degenerate_reversed_force"]
    # the sign of zero differs between the orders exactly as in the report
    assert math.copysign(1.0, fwd[0, 0]) == math.copysign(1.0, inv["degenerate_forward_force"][0])
    # hand-worked: canonical tangent (1, 0) -> fallback normal (0, 1); moving
    # up (+y) the approach (v - v_h).n > 0 flips it to (0, -1); the speed
    # along the normal is then -0.3, so the magnitude is k r + c 0.3
    f, _ = hull_forces(on, [[0.0, 0.3]], hull, ZERO, **kw)
    assert f[0] == pytest.approx([0.0, -(1000.0 * 0.05 + 10.0 * 0.3)], rel=1e-12)
    f, _ = hull_forces(on, [[0.0, -0.3]], hull[::-1].copy(), ZERO, **kw)
    assert f[0] == pytest.approx([0.0, 1000.0 * 0.05 + 10.0 * 0.3], rel=1e-12)
    # a vertical segment: canonical tangent (0, 1) -> fallback normal (-1, 0)
    vert = np.array([[0.0, 1.0], [0.0, -1.0]])
    f, _ = hull_forces(on, on, vert, ZERO, **kw)
    assert f[0] == pytest.approx([-50.0, 0.0], rel=1e-12)


def test_planted_mistakes_are_rejected():
    """Misreadings of the law the pins above must catch."""
    # (1) normal from the raw (not canonical) tangent: the reversed polyline
    # would push the other way
    hull = np.array([[1.0, 0.0], [-1.0, 0.0]])
    t = (hull[1] - hull[0]) / 2.0
    raw_normal = np.array([-t[1], t[0]])
    f, _ = hull_forces(np.zeros((1, 2)), np.zeros((1, 2)), hull, ZERO, particle_radius=0.05,
                       hull_radius=0.0, stiffness=1000.0)
    assert not np.allclose(f[0], 50.0 * raw_normal)
    # (2) distance to the nearest vertex instead of the nearest point: the
    # sagitta law would not hold at chord midpoints
    radius, rp, seg = 0.5, 0.05, 32
    overlap = 0.5 * rp
    ang = np.array([math.pi / seg])
    pos = (rp + radius - overlap) * np.stack([np.cos(ang), np.sin(ang)], axis=1)
    f, _ = hull_forces(pos, np.zeros_like(pos), bow(radius, seg), ZERO, particle_radius=rp,
                       hull_radius=0.0, stiffness=1000.0)
    vtx = bow(radius, seg)[:2]
    d_vertex = np.min(np.sqrt(((vtx - pos) ** 2).sum(axis=1)))
    wrong = 1000.0 * max(rp - d_vertex, 0.0)
    assert abs(float(np.sqrt(f[0] @ f[0])) - wrong) > 1e-3
    # (3) damping allowed to pull (no clamp at zero) would give a negative force
    f, _ = hull_forces([[0.0, 0.04]], [[0.0, 10.0]], [[-1.0, 0.0], [1.0, 0.0]], ZERO,
                       particle_radius=0.05, hull_radius=0.0, stiffness=100.0, damping=100.0)
    assert f[0, 1] != 100.0 * 0.01 - 100.0 * 10.0
