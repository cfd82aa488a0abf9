"""radtranbook_amd -- MI355X-native implicit Monte Carlo (Fleck-Cummings IMC) engine.

Drop-in modules for the reference's IMC particle path:

* ``radtranbook_amd.IMC2D``  -- gray 2-D xy / rz (IMC/IMC2D.py)

The particle kernels are hand-written HIP for gfx950 in ``csrc/``, built into
``lib/libimc_hip.so`` (``python -m radtranbook_amd.build``) and called through
the C-ABI of ``include/imc_hip.h``.
"""
__all__ = ["IMC2D"]
