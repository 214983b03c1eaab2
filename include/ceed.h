/* Umbrella header of the libCEED C API provided by this library (see ceed/ceed.h). */
#include "ceed/ceed.h"
