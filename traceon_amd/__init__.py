"""traceon_amd — Traceon's radially symmetric BEM path (fill, solve, field sums, axial
interpolation, RK45 tracing) on AMD Instinct MI355X (gfx950) HIP kernels.

Modules mirror the reference package's layout for this path:
``backend`` (ctypes binding of libtraceon_hip.so), ``excitation``, ``solver``, ``field``,
``tracing``; ``device`` holds the device-resident building blocks, ``mesh``/``configs``
the synthetic geometries, ``surrogate`` the learned field model.
"""
__version__ = "0.1.0"
