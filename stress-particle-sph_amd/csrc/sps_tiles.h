// Tile width and window capacities of the tiled pair kernels.
// tune_tiles.py rewrites this file with the fastest candidate it measured;
// a compiler definition (-DSPS_TILE=... and the others) overrides it.
// Measured (ms per step: FP64 on the 1 M slope; batched ensemble, mixed):
//   t256   (256, 24, 640, 192, 0): 4.342, 9.9385 (relative 1.0)
//   t256w4 (256, 24, 640, 192, 4): 4.3632, 9.9289 (relative 1.002)
//   t256s  (256, 24, 576, 160, 0): 4.3852, 12.2164 (relative 1.1142)
//   t256sw4 (256, 24, 576, 160, 4): 4.3775, 12.2173 (relative 1.1133)
#pragma once
#ifndef SPS_TILE
#define SPS_TILE 256  // receivers per workgroup
#endif
#ifndef SPS_TROWS
#define SPS_TROWS 24  // most window rows staged
#endif
#ifndef SPS_CAP
#define SPS_CAP 640  // staged partners of a main kind per window
#endif
#ifndef SPS_CAP_D
#define SPS_CAP_D 192  // staged dummies per window
#endif
#ifndef SPS_STAGE_WAVES
#define SPS_STAGE_WAVES 0  // waves per SIMD of the FP64 stage kernels (0: the compiler's choice)
#endif
