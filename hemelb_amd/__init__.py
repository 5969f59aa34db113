"""MI355X-native HemeLB lattice-Boltzmann timestep.

``libhemelb_amd.so`` (HIP kernels for gfx950 plus the C ABI declared in
``include/hemelb_amd.h`` and ``include/hemelb_geometry.h``) does all the work;
this package mirrors the reference's ``hemelb::lb`` names on the host.
"""
from . import _native  # noqa: F401
from .config import SimBuilder, SimConfig, read_xml, write_xml  # noqa: F401
from .geometry import Domain, Geometry, IoletPlane  # noqa: F401
from .lb import (D3Q15, D3Q19, LBGK, LBM, MRT, BouzidiFirdaousLallemand,  # noqa: F401
                 HaloPlan, InOutLetCosine, InOutLetParabolicVelocity, InOutLetWomersleyVelocity,
                 LaddIolet, LbmParameters, NashZerothOrderPressure, SimpleBounceBack,
                 Traits, UnitConverter)

__all__ = [
    "Domain", "Geometry", "IoletPlane", "D3Q15", "D3Q19", "LBGK", "MRT", "LBM",
    "SimpleBounceBack", "BouzidiFirdaousLallemand", "NashZerothOrderPressure",
    "LaddIolet", "InOutLetCosine", "InOutLetParabolicVelocity",
    "InOutLetWomersleyVelocity", "LbmParameters", "Traits", "UnitConverter",
    "SimConfig", "SimBuilder", "read_xml", "write_xml", "HaloPlan",
]
