/* ORACLE — TEST INFRASTRUCTURE ONLY: the double-precision gridder
 * (support <= 16, tiles of 16 cells; wgridder.rs:97-112, 1161-1172). */
#define REAL double
#define SUF _f64
#define SINGLE 0
#define WMAXT 16
#define LOG2TILE 4
#include "wg_impl.h"
