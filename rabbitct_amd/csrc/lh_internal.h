// lh_internal.h -- declarations shared by the device-layer translation units.
#ifndef LH_INTERNAL_H
#define LH_INTERNAL_H

#include <hip/hip_runtime.h>

#include "lolahip.h"

// Records the calling thread's error message and returns code.
int lh_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const LhConfig &lh_config(const LhContext *c);
hipStream_t lh_stream(const LhContext *c);

#endif
