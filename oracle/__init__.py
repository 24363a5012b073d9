"""FP64 NumPy oracle for the Shack-Hartmann hot path (test infrastructure only).

This package restates, line by line, the reference's CPU algorithm for the SH
forward chain, the ShSlopec centre of gravity, the CCD rebin and the LGS
convolution-kernel build.  It is the checker: only the test suite, the smoke
check and the CPU-baseline leg of the benchmark may import it.  The product
(``specula_amd``) never imports it and has no CPU fallback.
"""
