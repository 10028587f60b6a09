// the one-pass VALU level with Dk = 3, Dp = 2 fixed at compile time (fused_fixed.inc)
#define FSH_FDK 3
#define FSH_FDP 2
#include "fused_fixed.inc"
