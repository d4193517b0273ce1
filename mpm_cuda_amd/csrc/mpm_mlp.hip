// mpm_mlp.hip -- placeholder translation unit for the warm-start network.
#include "mpm_internal.h"
