"""MI355X-native Ensemble Monte Carlo particle step (ViennaEMC's hot path).

The compute path is the HIP library libemcgpu.so (C ABI in include/emcgpu.h);
:mod:`viennaemc_amd.emcgpu` binds it with ctypes and
:mod:`viennaemc_amd.bulk` mirrors the reference's bulk particle handler.
"""
