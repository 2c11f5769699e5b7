"""MI355X-native fragment-ion-index peptide search (GiCOPS hot path).

Host side of the C ABI in include/gicops_amd.h: `slm` builds indexes and
query batches, `search` runs the HIP search + e-value kernels.
"""
