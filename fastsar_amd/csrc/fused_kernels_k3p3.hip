// the one-pass VALU level with Dk = 3, Dp = 3 fixed at compile time (fused_fixed.inc)
#define FSH_FDK 3
#define FSH_FDP 3
#include "fused_fixed.inc"
