// Kernel instantiations: storage double, complex=false.
#define FKD_S double
#define FKD_CX false
#define FKD_FIND find_kset_dr
#include "kpm_inst.inc"
