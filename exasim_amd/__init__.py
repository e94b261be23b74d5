"""MI355X-native HDG Newton-GMRES path for Exasim inputs.

The solver itself lives in ``libexahip.so`` (C++ host code + HIP kernels for
gfx950) behind the C ABI declared in ``include/exasim_hdg.h``; ``exasim_amd.hdg``
is a thin ctypes mirror of that interface.  ``datain`` and ``cases`` write the
text2code binary input layout.
"""
