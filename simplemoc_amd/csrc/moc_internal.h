// moc_internal.h -- shared declarations of the simplemoc_amd library.
#ifndef MOC_INTERNAL_H
#define MOC_INTERNAL_H

#include <cstdarg>
#include <cstdint>

#include "../../include/simplemoc.h"

int moc_set_error(int code, const char *fmt, ...);

#endif
