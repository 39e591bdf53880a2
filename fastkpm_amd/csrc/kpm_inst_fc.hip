// Kernel instantiations: storage float, complex=true.
#define FKD_S float
#define FKD_CX true
#define FKD_FIND find_kset_fc
#include "kpm_inst.inc"
