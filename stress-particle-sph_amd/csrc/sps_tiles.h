// Tile width and window capacities of the tiled pair kernels.
// tune_tiles.py rewrites this file with the fastest candidate it measured;
// a compiler definition (-DSPS_TILE=... and the others) overrides it.
// Measured (ms per step: FP64 on the 1 M slope; batched ensemble, mixed):
//   t256   (256, 24, 640, 192, 0): 3.7794, 9.7972 (relative 1.0)
//   t256m  (256, 24, 608, 192, 0): 3.7552, 9.7305 (relative 0.9934)
#pragma once
#ifndef SPS_TILE
#define SPS_TILE 256  // receivers per workgroup
#endif
#ifndef SPS_TROWS
#define SPS_TROWS 24  // most window rows staged
#endif
#ifndef SPS_CAP
#define SPS_CAP 608  // staged partners of a main kind per window
#endif
#ifndef SPS_CAP_D
#define SPS_CAP_D 192  // staged dummies per window
#endif
#ifndef SPS_STAGE_WAVES
#define SPS_STAGE_WAVES 0  // waves per SIMD of the FP64 stage kernels (0: the compiler's choice)
#endif
