// sprgpu_internal.h -- shared declarations of the sprgpu library sources.
#ifndef SPRGPU_INTERNAL_H
#define SPRGPU_INTERNAL_H

#include "../../include/sprgpu.h"

// Records a message for sprg_last_error() and returns the code.
int sprg_fail(int code, const char* fmt, ...)
    __attribute__((format(printf, 2, 3)));

#endif
