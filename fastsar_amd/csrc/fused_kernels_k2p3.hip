// the one-pass VALU level with Dk = 2, Dp = 3 fixed at compile time (fused_fixed.inc)
#define FSH_FDK 2
#define FSH_FDP 3
#include "fused_fixed.inc"
