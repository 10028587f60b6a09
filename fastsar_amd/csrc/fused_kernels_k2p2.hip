// the one-pass VALU level with Dk = 2, Dp = 2 fixed at compile time (fused_fixed.inc)
#define FSH_FDK 2
#define FSH_FDP 2
#include "fused_fixed.inc"
