"""Exception types of the drop-in boundary.

isce3 raises ``isce3::except::*`` C++ exceptions that pybind11 translates by
their standard base class (cxx/isce3/except/Error.h:30-36): invalid_argument,
domain_error and length_error become ValueError, out_of_range becomes
IndexError, runtime_error becomes RuntimeError.  These subclasses keep those
bases so ``except ValueError`` style handlers written against isce3 still work.
"""


class InvalidArgument(ValueError):
    """isce3::except::InvalidArgument (std::invalid_argument -> ValueError)."""


class DomainError(ValueError):
    """isce3::except::DomainError (std::domain_error -> ValueError)."""


class LengthError(ValueError):
    """isce3::except::LengthError (std::length_error -> ValueError)."""


class OutOfRange(IndexError):
    """isce3::except::OutOfRange (std::out_of_range -> IndexError)."""


class RuntimeError_(RuntimeError):
    """isce3::except::RuntimeError (std::runtime_error -> RuntimeError)."""


class HipError(RuntimeError):
    """A HIP runtime call failed (the CUDA path raises isce3::cuda::except::CudaError)."""
