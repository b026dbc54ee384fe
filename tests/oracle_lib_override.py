"""pytest plugin (``-p tests.oracle_lib_override``): point the oracle's loader at the
shared library named by ORACLE_LIB instead of oracle/_build/liboracle.so.  Used by
tests/test_oracle_mutants.py to run the closed-form pins against an oracle built with
one planted mistake."""
import os

import oracle.pyoracle as po

po._LIB = os.environ["ORACLE_LIB"]
