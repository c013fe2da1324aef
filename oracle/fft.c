/* fft.c - instantiates the oracle FFT in float and double.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL float
#define SFX f32
#include "fft_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX f64
#include "fft_impl.h"
#undef REAL
#undef SFX

void orc_set_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
