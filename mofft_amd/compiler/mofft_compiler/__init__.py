"""gfx950 kernel generator for the MoFFT radix stages.

Replaces the reference's SME generator (compiler/mofft_compiler/emitter.py,
backends/arm_sme.py).  It emits:

* straight-line radix-r DFT codelets (``graph`` + ``codelets``), each an
  exact-arithmetic rewrite of the r x r DFT-matrix contraction
  (direct, like-terms pairing, or a Cooley-Tukey factorisation that
  generalises the reference's even-odd rewrite, rewrites.py:43-95);
* the kernel registry (``registry``): which LDS-resident pass kernels are
  compiled for gfx950, with their sub-stage chains and tile shapes;
* HIP sources instantiating the pass kernel template for every registry
  entry, plus a host-side C table of the registry (``emit``).
"""
__all__ = ["graph", "codelets", "registry", "emit"]
