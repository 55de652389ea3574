"""helios_amd: MI355X-native closed-shell PPP CCSD (helios method ``r-ccsd``).

The compute path is libhelios_ccsd.so (HIP, gfx950), reached through its C
ABI (include/helios_ccsd.h): from Fortran through the iso_c_binding module in
helios_amd/src/modules, from Python through ``helios_amd.capi``.
"""
