/* FP32 instantiation of the oracle propagators: the CPU baseline of the
 * stencil (same algorithm and operation order as the FP64 oracle). */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <xmmintrin.h>
#include "fwi_oracle.h"
/* flush-to-zero and denormals-are-zero on every worker thread, as a tuned
 * FP32 CPU stencil runs (denormal tails otherwise slow the sweep). */
#define OR_THREAD_SETUP() _mm_setcsr(_mm_getcsr() | 0x8040)
#define REAL float
#define SFX(name) name##_f32
#include "oracle_core.inc"
