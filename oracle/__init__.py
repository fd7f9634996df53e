"""The oracle: a CPU restatement of HydroC99's split Godunov/MUSCL step.

TEST INFRASTRUCTURE ONLY.  ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this package, as the reference check.  The
product (``hydro_amd``) never imports it, and nothing measured runs through it.

Two implementations sit behind the same Python surface:

* ``Port``  -- ``oracle/_port/libhydro_oracle.so``, built from
  ``oracle/hydro_oracle.c`` (this repository's own C99 restatement);
* ``Ref``   -- ``oracle/_ref/libhydroref.so``, the reference HydroC99_2DMpi
  sources compiled in place by ``oracle/Makefile`` plus ``ref_harness.c``.
  Present only where it was built (it travels with the tree, the reference
  sources do not).
"""
from .bindings import (  # noqa: F401
    ORACLE_DIR, Params, Port, Ref, build, deck_text, port_available, ref_available,
    write_deck,
)
