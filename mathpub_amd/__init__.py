"""mathpub_amd: MI355X-native engine for the Wieferich (A001220) sieve -> Fermat-quotient path.

The compute path is libwf.so (hand-written HIP for gfx950, C ABI in include/wf.h); this
package is its Python mirror (mathpub_amd.wf).  There is no CPU fallback: every call goes
to the GPU through the C ABI and fails loudly when the library or the GPU is missing.
"""
from .wf import Engine, WFError, load  # noqa: F401
