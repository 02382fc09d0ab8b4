// Mixed-precision reciprocal-space sum (split-FP16 MFMA): placeholder until
// the MFMA path is built; it refuses to run rather than falling back.
#include "mgpu_internal.h"

void mgpu::run_ewald_mixed(mgpu_engine *, double *) {
   throw mgpu::StatusError{MGPU_E_STATE, "mixed-precision Ewald path not built"};
}
