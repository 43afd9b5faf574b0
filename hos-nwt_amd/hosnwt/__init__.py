"""Host side of the MI355X HOS-NWT time-stepping core.

``lib``   : ctypes binding of the C-ABI (include/hosnwt.h)
``setup`` : the driver's setup restated for the host (nondimensional
            parameters, wave components, paddle transfer functions, beach,
            RK4 stability bound) - HOS-NWT.f90:95-255 and sources/wave
``tank``  : ``Tank``, the driver loop around the GPU stepper
            (HOS-NWT.f90:446-581)
"""
from . import lib  # noqa: F401
