// render.hip -- render kernels (Path B).
#include "bt_internal.h"
