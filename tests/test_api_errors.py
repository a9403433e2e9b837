"""Argument checks of the drop-in backproject, which raise the same exception
types as the reference binding (pybind focus/Backproject.cpp:29-151,
cuda/focus/Backproject.cpp:87-89, DryTroposphereModel.cpp:20-33,
Backproject.cpp:88-92) before any device work."""
import numpy as np
import pytest

from isce3_amd import core, errors
from isce3_amd.container import RadarGeometry
from isce3_amd.focus import backproject
from isce3_amd.sim import make_scene


@pytest.fixture(scope="module")
def setup():
    sc = make_scene("tiny")
    out = np.zeros(sc.out_grid.shape, np.complex64)
    raw = np.zeros(sc.in_grid.shape, np.complex64)
    return sc, out, raw


def call(sc, out, raw, **kw):
    args = dict(dry_tropo_model="nodelay")
    args.update(kw)
    return backproject(out, sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                       sc.kernel, **args)


def test_shapes(setup):
    sc, out, raw = setup
    with pytest.raises(ValueError, match="output array must be 2-D"):
        call(sc, out.ravel(), raw)
    with pytest.raises(ValueError, match="output array shape must match"):
        call(sc, out[:-1], raw)
    with pytest.raises(ValueError, match="input signal data must be 2-D"):
        call(sc, out, raw.ravel())
    with pytest.raises(ValueError, match="input signal data shape must match"):
        call(sc, out, raw[:, :-1])
    with pytest.raises(ValueError, match="height array shape must match"):
        call(sc, out, raw, height=np.zeros((3, 3), np.float32))


def test_output_must_be_writable_complex64_c_contiguous(setup):
    sc, out, raw = setup
    with pytest.raises(errors.InvalidArgument):
        call(sc, out.astype(np.complex128), raw)
    with pytest.raises(errors.InvalidArgument):
        call(sc, np.asfortranarray(out), raw)
    with pytest.raises(errors.InvalidArgument):
        call(sc, out, raw, height=np.zeros(sc.out_grid.shape, np.float64))


def test_tropo_and_param_dicts(setup):
    sc, out, raw = setup
    with pytest.raises(ValueError, match="expected one of"):
        call(sc, out, raw, dry_tropo_model="saastamoinen")
    with pytest.raises(ValueError, match="unexpected rdr2geo_bracket keyword: tol"):
        call(sc, out, raw, rdr2geo_params={"tol": 1.0})
    with pytest.raises(ValueError, match="unexpected geo2rdr_bracket keyword: threshold"):
        call(sc, out, raw, geo2rdr_params={"threshold": 1.0})


def test_batch_must_be_positive(setup):
    sc, out, raw = setup
    with pytest.raises(errors.DomainError):
        call(sc, out, raw, batch=0)


def test_epoch_mismatch(setup):
    sc, out, raw = setup
    orb = sc.orbit.copy()
    orb.update_reference_epoch(core.DateTime("2025-01-01T00:00:01"))
    og = RadarGeometry(sc.out_grid, orb, sc.doppler)
    with pytest.raises(RuntimeError, match="reference epoch"):
        backproject(out, og, raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds, sc.kernel, "nodelay")


def test_kernel_and_precision(setup):
    sc, out, raw = setup
    with pytest.raises(errors.InvalidArgument):
        backproject(out, sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds,
                    core.KnabKernel(8, 0.8), "nodelay")
    with pytest.raises(errors.InvalidArgument):
        call(sc, out, raw, precision="half")


def test_parsers_accept_reference_keys():
    from isce3_amd.focus import parse_geo2rdr_params, parse_rdr2geo_params
    p = parse_rdr2geo_params({"tol_height": 1e-4, "look_min": 0.1, "look_max": 1.0})
    assert (p.tol_height, p.look_min, p.look_max) == (1e-4, 0.1, 1.0)
    q = parse_geo2rdr_params({"tol_aztime": 1e-6, "time_start": None, "time_end": 5.0})
    assert q.tol_aztime == 1e-6 and q.have_time_start == 0 and q.have_time_end == 1
    assert q.time_end == 5.0


def test_block_list_checks(setup):
    """backproject_blocks rejects inconsistent block lists before device work."""
    from isce3_amd.focus import backproject_blocks
    from isce3_amd.sim import straight_orbit
    sc, out, raw = setup
    og = sc.out_geometry()
    args = (raw, sc.in_geometry(), sc.dem, sc.fc, sc.ds, sc.kernel, "nodelay")
    assert backproject_blocks([], [], *args) == []
    with pytest.raises(ValueError, match="one output geometry per output array"):
        backproject_blocks([out], [og, og], *args)
    with pytest.raises(ValueError, match="one height entry"):
        backproject_blocks([out], [og], *args, heights=[None, None])
    with pytest.raises(ValueError, match="output array shape must match"):
        backproject_blocks([out[:-1]], [og], *args)
    other = RadarGeometry(sc.out_grid, straight_orbit(60.0, 10.0, sc.orbit.reference_epoch),
                          sc.doppler)
    with pytest.raises(ValueError, match="share one orbit"):
        backproject_blocks([out, out], [og, other], *args)
    with pytest.raises(ValueError, match="precision"):
        backproject_blocks([out], [og], *args, precision="fp16")
