"""MI355X-native RI-V Coulomb/exchange (J/K) build — a drop-in for the vj/vk
builders of REST's SCF (src/scf_io/mod.rs:2386-2834).

The product is the C-ABI library rest_amd/lib/librest_jk.so (include/rest_jk.h);
rest_amd.jk is its Python host binding.
"""
from .jk import (DeviceRI, RJKError, PREC_FP32, PREC_FP64, STORE_FP32, STORE_FP64,  # noqa: F401
                 vj_upper_with_rimatr_v_sync, vj_upper_with_ri_v_sync, vk_upper_with_ri_v_sync,
                 vk_upper_with_rimatr_use_dm_only_sync, vk_upper_with_ri_v_use_dm_only_sync)
