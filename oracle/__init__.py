"""CPU FP64 oracle for the rskelf path (test infrastructure only).

This package restates, in NumPy/SciPy, the reference algorithms that the
HIP library replaces: the adaptive tree (core/hypoct.m), the interpolative
decomposition with Gu-Eisenstat refinement (core/id.m), the recursive
skeletonization factorization (rskelf/rskelf.m) and its solve/apply sweeps
(rskelf/sv/*.m, rskelf/mv/*.m), plus the problem definitions of
rskelf/test/ie_circle.m and rskelf/test/ie_ellipse.m.

It is the checker: only tests, the smoke entry point and the CPU-baseline
leg of the benchmark may import it.  The product library never calls it.
"""
