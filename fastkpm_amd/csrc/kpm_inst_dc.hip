// Kernel instantiations: storage double, complex=true.
#define FKD_S double
#define FKD_CX true
#define FKD_FIND find_kset_dc
#include "kpm_inst.inc"
