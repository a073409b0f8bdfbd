"""tabipb -- MI355X host package for the treecode-accelerated boundary-integral
Poisson-Boltzmann solve.

* meshio  : MSMS / PQR / usrdata.in formats and the icosphere generator
* driver  : the reference driver's discretisation and set-up, on the host
* hip     : ctypes binding of libtabi_hip.so (the GPU path)
"""
from . import meshio  # noqa: F401

__all__ = ["meshio", "driver", "hip"]
