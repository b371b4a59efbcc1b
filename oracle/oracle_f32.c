/* FP32 instantiation of the oracle propagators: the CPU baseline of the
 * stencil (same algorithm and operation order as the FP64 oracle). */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <xmmintrin.h>
#include "fwi_oracle.h"
/* flush-to-zero and denormals-are-zero on every worker thread for the pass,
 * as a tuned FP32 CPU stencil runs (denormal tails otherwise slow the
 * sweep); each thread's previous MXCSR is put back at the end of the pass,
 * so the caller's later arithmetic keeps its subnormals. */
static __thread unsigned or_csr_saved;
#define OR_THREAD_SETUP() (or_csr_saved = _mm_getcsr(), _mm_setcsr(or_csr_saved | 0x8040))
#define OR_THREAD_RESTORE() _mm_setcsr(or_csr_saved)
#define REAL float
#define SFX(name) name##_f32
#include "oracle_core.inc"
