/* Driver of the drop-in link test (tests/test_dropin.py): defines what a
 * Meraxes build would (run_globals, mlog, mlog_error), then calls the bubble
 * finder's GPU entry point without init_CUDA having run, so the library
 * returns MRX_EINVAL (no context) and the shim reports it the reference's
 * way: mlog_error, then ABORT - which exits here with status 3. */
#include <stdarg.h>
#include <stdbool.h>
#include <stdio.h>

#include "meraxes.h"

run_globals_t run_globals;

void mlog(const char* fmt, int type, ...)
{
  (void)type;
  va_list ap;
  va_start(ap, type);
  vfprintf(stdout, fmt, ap);
  va_end(ap);
}

void mlog_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fputc('\n', stderr);
}

void stub_abort(int sigterm)
{
  fprintf(stderr, "ABORT(%d)\n", sigterm);
  exit(3);
}

void _find_HII_bubbles_gpu(const int snapshot, const bool flag_write_validation_output);
void _ComputeTs_gpu(int snapshot);
void ComputeBrightnessTemperatureBox_gpu(int snapshot);
void construct_baryon_grids_gpu(int snapshot, int local_ngals);
void init_CUDA(void);

int main(void)
{
  static double zz[2] = {8.0, 8.0};
  static ptrdiff_t nix[1] = {32};
  run_globals.ZZ = zz;
  run_globals.params.ReionGridDim = 32;
  run_globals.reion_grids.slab_nix = nix;
  _find_HII_bubbles_gpu(0, false);
  return 0; /* not reached: the library refuses a call without a context */
}
