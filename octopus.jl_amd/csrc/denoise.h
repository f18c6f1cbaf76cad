// denoise.h -- the residual denoiser as the collision path uses it (k_denoise.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "octopic.h"
