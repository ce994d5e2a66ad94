"""phylocaml_amd -- MI355X (gfx950) implementation of phylocaml's two
site-pattern hot paths: Fitch parsimony (NonAdditive_c) and Felsenstein
pruning (Likelihood_c over MlModel).

The compute lives in libphylo_hip.so (hand-written HIP kernels behind the C
ABI of include/phylo_hip.h); this package is the host-side mirror of the
reference's module interfaces.  Importing it requires the built library.
"""
from . import _lib  # noqa: F401  (raises ImportError if the library is missing)
from . import device, likelihood, mlmodel, nonadditive, synth  # noqa: F401

__all__ = ["device", "likelihood", "mlmodel", "nonadditive", "synth"]
