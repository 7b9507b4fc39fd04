/* ORACLE — TEST INFRASTRUCTURE ONLY: the single-precision gridder
 * (ducc0's f32 kernels, support <= 8, tiles of 32 cells; wgridder.rs:80-95,
 * 1161-1172). */
#define REAL float
#define SUF _f32
#define SINGLE 1
#define WMAXT 8
#define LOG2TILE 5
#include "wg_impl.h"
