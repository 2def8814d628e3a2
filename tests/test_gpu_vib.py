"""The harmonic-analysis readout (hh_vib_*, vib.rs:52-552) run as part of the GPU suite.

The readout is host code inside the shipped libhesship.so; the same tests as tests/test_vib.py
(reference NH3 fixtures of tests/test_vib.rs and oracle parity) run here under the gpu marker so
that the library built for and loaded on the GPU machine is the one checked."""
import pytest

from tests.test_vib import (test_library_golden_rotation_constants, test_library_matches_oracle,  # noqa: F401
                            test_library_rejects_bad_input, test_library_tr_space)

pytestmark = pytest.mark.gpu
