"""Enumerations of the kernel-creation API.

Same member names as lbmpy's ``src/lbmpy/enums.py`` (``Stencil`` :4-40,
``Method`` :42-134, ``CollisionSpace`` :136-175, ``ForceModel`` :178-223,
``SubgridScaleModel`` :226-238) so that user code written against lbmpy
selects the same configurations here.  ``Target`` mirrors pystencils'
``Target`` (CPU/GPU), which lbmpy takes through ``CreateKernelConfig``.
"""
from enum import Enum, auto


class Stencil(Enum):
    D2Q9 = auto()
    D2V17 = auto()
    D2V37 = auto()
    D3Q7 = auto()
    D3Q15 = auto()
    D3Q19 = auto()
    D3Q27 = auto()


class Method(Enum):
    SRT = auto()
    TRT = auto()
    MRT_RAW = auto()
    MRT = auto()
    CENTRAL_MOMENT = auto()
    TRT_KBC_N1 = auto()
    TRT_KBC_N2 = auto()
    TRT_KBC_N3 = auto()
    TRT_KBC_N4 = auto()
    CUMULANT = auto()
    MONOMIAL_CUMULANT = auto()


class CollisionSpace(Enum):
    POPULATIONS = auto()
    RAW_MOMENTS = auto()
    CENTRAL_MOMENTS = auto()
    CUMULANTS = auto()

    def compatible(self, method):
        """Which methods may be derived in this collision space (lbmpy ``enums.py:165-175``)."""
        compat = {
            CollisionSpace.POPULATIONS: {Method.SRT, Method.TRT, Method.MRT_RAW, Method.MRT,
                                         Method.TRT_KBC_N1, Method.TRT_KBC_N2, Method.TRT_KBC_N3,
                                         Method.TRT_KBC_N4},
            CollisionSpace.RAW_MOMENTS: {Method.SRT, Method.TRT, Method.MRT_RAW, Method.MRT},
            CollisionSpace.CENTRAL_MOMENTS: {Method.CENTRAL_MOMENT},
            CollisionSpace.CUMULANTS: {Method.MONOMIAL_CUMULANT, Method.CUMULANT},
        }
        return method in compat[self]


class ForceModel(Enum):
    SIMPLE = auto()
    LUO = auto()
    GUO = auto()
    BUICK = auto()
    SILVA = auto()
    EDM = auto()
    KUPERSHTOKH = auto()
    HE = auto()
    SHANCHEN = auto()
    CENTRALMOMENT = auto()


class SubgridScaleModel(Enum):
    SMAGORINSKY = auto()
    QR = auto()


class Target(Enum):
    """Where a kernel runs (pystencils ``Target``)."""
    CPU = auto()
    GPU = auto()
