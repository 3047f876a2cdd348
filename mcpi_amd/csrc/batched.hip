// Batched multi-instance Lanczos (filled in below the single-instance path).
#include <hip/hip_runtime.h>

#include "mcpi_internal.h"
