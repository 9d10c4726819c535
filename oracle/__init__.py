"""CPU oracle for newt's step: test infrastructure only (see newt_ref.hpp)."""
