"""gsFileData XML input and point / CSV output (xmlio.py) on the CPU: the reader against the
programmatic builders, a round trip through the writer, the reference's own problem files when
they are present, the expression evaluator, and the output formats."""
import os

import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g
from gsstructuralanalysis_amd import xmlio

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/filedata/pde"


def _same_patch(a, b, tol=1e-13):
    assert (a.p, a.nx, a.ny) == (b.p, b.nx, b.ny)
    assert np.abs(a.ctrl - b.ctrl).max() <= tol * max(1.0, np.abs(b.ctrl).max())
    assert np.array_equal(a.fixed, b.fixed)
    assert (a.t, a.E, a.nu, a.rho) == (b.t, b.E, b.nu, b.rho)
    assert np.allclose(a.point_xi, b.point_xi, rtol=0, atol=1e-15) and np.allclose(a.point_F, b.point_F, rtol=1e-15)
    assert tuple(a.fsurf) == tuple(b.fsurf)
    assert a.pressure == b.pressure


@pytest.mark.parametrize("refine", [2, 5])
def test_roof_file_reproduces_the_roof_builder(refine):
    """The quadratic one-patch roof file, elevated once and refined r times, is the cubic
    2^r x 2^r roof of the arc-length benchmark (the refined space contains the quadratic surface)."""
    x = xmlio.read_problem(os.path.join(HERE, "golden", "roof.xml"), elevate=1, refine=refine)
    _same_patch(x.patch, g.roof_problem(p=3, n=2 ** refine))
    assert x.ref_points.shape == (2, 2) and np.allclose(x.ref_points.T, [[0.5, 0.5], [0.0, 0.5]])
    kw = xmlio.alm_kwargs(x.options[93])
    assert kw["method"] == 0 and kw["tol_f"] == 1e-8 and kw["scaling"] == 0.0 and x.options[93]["Length"] == 2e-3


def _pressurised_roof():
    pt = g.roof_problem(p=3, n=4, P=0.0)
    pt.pressure = -2.5
    return pt


@pytest.mark.parametrize("build", [lambda: g.clamped_plate_problem(p=2, n=6, q=3.0),
                                   lambda: g.compressed_strip_problem(nx=8),
                                   lambda: g.roof_problem(p=3, n=4), _pressurised_roof])
def test_write_then_read_round_trip(build, tmp_path):
    pt = build()
    f = str(tmp_path / "problem.xml")
    xmlio.write_problem(pt, f, ref_points=[[0.5], [0.5]], options={91: {"tol": 1e-10, "maxIt": 25}})
    x = xmlio.read_problem(f)
    _same_patch(x.patch, pt)
    assert x.options[91] == {"tol": 1e-10, "maxIt": 25}


def test_unrepresentable_content_is_refused(tmp_path):
    pt = g.clamped_plate_problem(p=2, n=2)
    f = str(tmp_path / "p.xml")
    xmlio.write_problem(pt, f)
    text = open(f).read()
    for bad, msg in ((text.replace('type="Linear3"', 'type="CompressibleNH3"'), "St. Venant"),
                     (text.replace('type="TensorBSpline2"', 'type="TensorNurbs2"'), "NURBS"),
                     (text.replace('<bc type="Clamped"', '<bc type="Collapsed"', 1), "Collapsed")):
        open(f, "w").write(bad)
        with pytest.raises(xmlio.XmlError, match=msg):
            xmlio.read_problem(f)


def test_constant_expressions():
    assert xmlio.constant_expr("2*2*(15.8114570e4+6.21485502e4)*(1+0.5)") == pytest.approx(
        2 * 2 * (15.8114570e4 + 6.21485502e4) * 1.5, rel=1e-15)
    assert xmlio.constant_expr("2^3") == 8.0
    assert xmlio.constant_expr("-sqrt(4) + pi") == pytest.approx(np.pi - 2)
    for bad in ("x + 1", "__import__('os')", "(1).real", "[1]"):
        with pytest.raises(xmlio.XmlError):
            xmlio.constant_expr(bad)


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's data files are not present")
def test_reference_problem_files():
    """The reference's own files: the SvK single-patch ones read (degree elevation and refinement
    as the drivers do), the others are refused with the reason."""
    x = xmlio.read_problem(os.path.join(REF, "cornerLoads_1p.xml"), elevate=2, refine=3)
    pt = x.patch
    assert (pt.p, pt.nx, pt.t, pt.E, pt.nu) == (3, 8, 50e-6, 2.96e9, 0.34)
    assert np.allclose(pt.ctrl[-1], [0.2, 0.2, 0.0]) and tuple(pt.fsurf) == (0.0, 0.0, 1e-2)
    assert np.all(pt.fixed[:3] == 1) and pt.nfree == pt.ndof - 3  # corner value on corner 1
    assert pt.point_xi.shape == (4, 2) and x.ref_points.shape == (2, 3)
    for name, msg in (("shear_1p.xml", "Collapsed"), ("pillow_1p.xml", "St. Venant"),
                      ("tension_1p.xml", "St. Venant"), ("kirchhoff_shell_scordelis.xml", "geometry")):
        with pytest.raises(xmlio.XmlError, match=msg):
            xmlio.read_problem(os.path.join(REF, name), elevate=1, refine=2)


def test_point_evaluation():
    pt = g.roof_problem(p=3, n=4)
    params = np.array([[0.0, 0.5, 1.0, 0.3], [0.0, 0.5, 1.0, 0.7]])
    # the surface itself (the roof's parabolic section, exactly reproduced)
    X = xmlio.geometry_at(pt, params)
    s, h = 2.5 * np.sin(0.1), 2.5 * np.sin(0.1) * np.tan(0.1)
    assert np.allclose(X, [0.5 * params[0], -2 * s * params[1], 2 * h * params[1] * (1 - params[1])], atol=1e-14)
    # a displacement field that the fixed dofs allow: a uniform x translation of the free dofs is
    # the field itself wherever every control point in the support is free
    full = np.zeros((pt.mx * pt.my, 3))
    full[:, 0] = 1.0
    u = pt.to_free(full.ravel())
    d = xmlio.displacement_at(pt, u, [[0.5], [0.5]])
    assert np.allclose(d[:, 0], [1.0, 0.0, 0.0], rtol=0, atol=1e-14)


def test_csv_formats(tmp_path):
    f = str(tmp_path / "out.csv")
    w = xmlio.StructuralAnalysisOutput(f, [[0.5, 1.0], [0.0, 0.5]])
    w.init(["x", "y", "z"], ["load"])
    w.add(np.array([[1.0, 4.0], [2.0, 5.0], [3.0, 6.123456789]]), [0.25])
    lines = open(f).read().splitlines()
    assert lines[0] == "point0-x,point0-y,point0-z,point1-x,point1-y,point1-z,load,"
    assert lines[1] == "1,2,3,4,5,6.1235,0.25,"
    pt = g.roof_problem(p=3, n=4)
    f2 = str(tmp_path / "steps.csv")
    U = np.zeros((2, pt.nfree))
    xmlio.write_step_csv(f2, pt, [[0.5], [0.5]], [0.0, 0.0], U, [0.1, 0.2], [1.0, -2.0])
    lines = open(f2).read().splitlines()
    assert lines[0] == "Deformation norm,point 0 - x,point 0 - y,point 0 - z,Lambda,Indicator"
    assert lines[2] == "0,0,0,0,0.2,-2,"
