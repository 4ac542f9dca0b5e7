"""Extension modules of the time-domain 2D propagator.

``td2d_cuda``, ``td2d_forward_cuda`` and ``td2d_exa_cuda`` carry the names and
the ``forward``/``backward`` argument lists of the reference's extension
modules (torchwi/propagator/td2d/wrapper_cuda.cpp, wrapper_forward_cuda.cpp,
wrapper_exa_cuda.cpp), so torchwi.operator.TimeFunction's backend selection
works unchanged; they call the gfx950 C-ABI library.
"""
