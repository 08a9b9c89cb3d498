"""MI355X-native Ensemble Monte Carlo particle step (ViennaEMC's hot path).

The compute path is the HIP library libemcgpu.so (C ABI in include/emcgpu.h);
:mod:`viennaemc_amd.emcgpu` binds it with ctypes; the C++17 host API that
mirrors the reference's particle type, bulk handler and emcSimulation is
header-only under viennaemc_amd/include/emcgpu/.
"""
