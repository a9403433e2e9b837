"""Range-compression boundary checks that need no GPU: the argument rules of
cxx/isce3/focus/RangeComp.cpp:39-79 and pybind focus/RangeComp.cpp:55-87,
raised before any device work, and the chirp generator against the oracle."""
import numpy as np
import pytest

import rangecomp as rco
from isce3_amd import errors
from isce3_amd.rangecomp import Mode, RangeComp, form_linear_chirp


def test_constructor_argument_rules():
    chirp = np.ones(5, np.complex64)
    with pytest.raises(errors.DomainError):
        RangeComp(chirp, 0)
    with pytest.raises(errors.DomainError):
        RangeComp(chirp, 10, maxbatch=0)
    with pytest.raises(RuntimeError):
        RangeComp(chirp, 10, mode=7)
    with pytest.raises(errors.InvalidArgument):
        # a chirp too long for overlap-save blocks and a convolution > 2^19 samples
        RangeComp(np.ones(10000, np.complex64), 520000)
    with pytest.raises(errors.InvalidArgument):
        RangeComp(np.zeros(0, np.complex64), 10)


def test_chirp_generator_matches_oracle():
    for args in [(1e12, 20e-6, 24e6), (100.0, 2.0, 2400.0, 5.0, 7.5, 0.5),
                 (-3e11, 11e-6, 40e6, 1e6)]:
        a = form_linear_chirp(*args)
        b = rco.form_linear_chirp(*args)
        assert a.shape == b.shape and len(a) % 2 == 1
        np.testing.assert_array_equal(a, b)
    with pytest.raises(errors.DomainError):
        form_linear_chirp(1.0, 0.0, 1.0)
    with pytest.raises(errors.DomainError):
        form_linear_chirp(1.0, 1.0, -1.0)
    with pytest.raises(errors.DomainError):
        form_linear_chirp(1.0, 1.0, 1.0, amplitude=0.0)


def test_mode_values_match_the_reference_enum():
    assert [Mode.Full, Mode.Valid, Mode.Same] == [0, 1, 2]
    assert RangeComp.Mode is Mode
