// Kernel instantiations: storage float, complex=false.
#define FKD_S float
#define FKD_CX false
#define FKD_FIND find_kset_fr
#include "kpm_inst.inc"
