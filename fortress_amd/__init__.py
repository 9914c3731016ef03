"""fortress_amd: MI355X (gfx950) split-step Fourier propagator for the coupled
Gross-Pitaevskii equations of spin-orbit-coupled spin-1/spin-2 condensates.

The product is the HIP library ``fortress_amd/libfgpe.so`` (C ABI declared in
``include/fgpe.h``) and the Fortran drop-in host ``fortress_amd/host/cgpe.x``.
This Python package is a thin ctypes binding over the same C ABI.
"""
from .abi import FgpeInput, FgpeParams  # noqa: F401
