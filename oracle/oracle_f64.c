/* FP64 instantiation of the oracle propagators (test infrastructure only). */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "fwi_oracle.h"
#define REAL double
#define SFX(name) name##_f64
#include "oracle_core.inc"
