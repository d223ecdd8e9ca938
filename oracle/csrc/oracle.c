/* C restatement of the reference SIMP hot path (TEST INFRASTRUCTURE):
 * checker for the large grids and the timed CPU baseline. Never linked into
 * the product. The body is written once (oracle_body.h) and instantiated in
 * FP64 (_f64) and FP32 (_f32, the reference's own precision). */
#include <math.h>
#include <omp.h>
#include <stdlib.h>

#define REAL double
#define SUF(x) x##_f64
#include "oracle_body.h"
#undef REAL
#undef SUF

#define REAL float
#define SUF(x) x##_f32
#include "oracle_body.h"
#undef REAL
#undef SUF

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }
