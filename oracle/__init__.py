"""CPU checkers for the TCP_allpair hot path -- TEST INFRASTRUCTURE ONLY.

`oracle.cpu` wraps liboracle.so, the C restatement of the reference's
force.f90 / vverlet.f90 path.  `oracle.refharness` drives _ref/ref_harness, the
reference's own Fortran routines compiled from their sources.  Only the tests,
the smoke check and the CPU-baseline leg of the benchmark import this package;
the product library never does.
"""
