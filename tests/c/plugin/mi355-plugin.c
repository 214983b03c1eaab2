/*
 * Runtime registration of /gpu/hip/mi355 into a libCEED that was built
 * without it: CeedRegister is public (include/ceed/backend.h), so a
 * constructor can add the backend before the application's first CeedInit.
 */
#include <ceed/backend.h>

int CeedInit_Mi355(const char *resource, Ceed ceed);

__attribute__((constructor)) static void mi355_plugin_register(void) { CeedRegister("/gpu/hip/mi355", CeedInit_Mi355, 20); }
